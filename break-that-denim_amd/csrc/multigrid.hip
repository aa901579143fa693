// Galerkin multigrid on the half-stencil operators: connectivity-masked prolongation weights,
// the Galerkin triple product P^T A P, residual / restriction / prolongation, and the
// colour-ordered block-tridiagonal yarn-line smoother.
//
// Restates buildProlongation / galerkin / vcycle / smoothLines / solvePenta
// (src/simulation/reference/newtonMultigrid.ts:287-528, 647-733) and lattice.ts:13-41, 86-106.
// The smoother solves each line exactly like solvePenta's banded elimination, but by odd-even
// block cyclic reduction with the whole line resident in LDS: one workgroup per line, one row per
// thread, log2(len) reduction levels, every pivot shifted by PIVOT_SHIFT of its scale.
#include <stdexcept>

#include "bd_internal.h"

namespace bd {

namespace {

__device__ __forceinline__ void blk_load(const float* op, int Np, int base, int row, float o[6]) {
#pragma unroll
  for (int q = 0; q < 6; q++) o[q] = op[(base + q) * Np + row];
}

__device__ __forceinline__ void blk_scalar(float s, float o[6]) {
  o[0] = s;
  o[1] = 0.f;
  o[2] = 0.f;
  o[3] = s;
  o[4] = 0.f;
  o[5] = s;
}

// every off-centre coupling (di, dj, neighbour m, symmetric block) of row n of a level operator
template <bool FINE, class Fn>
__device__ __forceinline__ void for_couplings(const float* op, int Np, int n, int i, int j, int nx, int ny, Fn&& fn) {
  float b[6];
  if (FINE) {
    const bool e = i + 1 < nx, w = i > 0, no = j + 1 < ny, so = j > 0;
    if (e) { blk_load(op, Np, F_E, n, b); fn(1, 0, n + 1, b); }
    if (w) { blk_load(op, Np, F_E, n - 1, b); fn(-1, 0, n - 1, b); }
    if (no) { blk_load(op, Np, F_N, n, b); fn(0, 1, n + nx, b); }
    if (so) { blk_load(op, Np, F_N, n - nx, b); fn(0, -1, n - nx, b); }
    if (e && no) { blk_load(op, Np, F_NE, n, b); fn(1, 1, n + nx + 1, b); }
    if (w && so) { blk_load(op, Np, F_NE, n - nx - 1, b); fn(-1, -1, n - nx - 1, b); }
    if (w && no) { blk_load(op, Np, F_NW, n, b); fn(-1, 1, n + nx - 1, b); }
    if (e && so) { blk_load(op, Np, F_NW, n - nx + 1, b); fn(1, -1, n - nx + 1, b); }
    if (i + 2 < nx) { blk_scalar(op[F_EE * Np + n], b); fn(2, 0, n + 2, b); }
    if (i >= 2) { blk_scalar(op[F_EE * Np + n - 2], b); fn(-2, 0, n - 2, b); }
    if (j + 2 < ny) { blk_scalar(op[F_NN * Np + n], b); fn(0, 2, n + 2 * nx, b); }
    if (j >= 2) { blk_scalar(op[F_NN * Np + n - 2 * nx], b); fn(0, -2, n - 2 * nx, b); }
  } else {
#pragma unroll
    for (int k = 1; k < kCoarseBlocks; k++) {
      const int di = coarse_di(k), dj = coarse_dj(k);
      if (i + di >= 0 && i + di < nx && j + dj < ny) {
        int m = n + dj * nx + di;
        blk_load(op, Np, 6 * k, n, b);
        fn(di, dj, m, b);
      }
      if (i - di >= 0 && i - di < nx && j - dj >= 0) {
        int m = n - dj * nx - di;
        blk_load(op, Np, 6 * k, m, b);
        fn(-di, -dj, m, b);
      }
    }
  }
}

struct LvArgs {
  const float* op;
  const uint8_t* fixed;
  int N, Np;
  Geo geo;
};

__device__ __forceinline__ void node_ij(const Geo& g, int n, int& l, int& i, int& j) {
  l = geo_layer(g, n);
  int r = n - g.off[l];
  j = r / g.nx[l];
  i = r - j * g.nx[l];
}

// y = A x on free rows (fixed rows give 0 here: the callers want residual-style masking)
template <bool FINE>
__device__ __forceinline__ void apply_row(const LvArgs& a, const float* x, int n, float y[3]) {
  int l, i, j;
  node_ij(a.geo, n, l, i, j);
  const int Np = a.Np;
  float c[6], xv[3] = {x[n], x[Np + n], x[2 * Np + n]};
  blk_load(a.op, Np, 0, n, c);
  sym_mul(c, xv, y);
  for_couplings<FINE>(a.op, Np, n, i, j, a.geo.nx[l], a.geo.ny[l], [&](int, int, int m, const float* b) {
    float xm[3] = {x[m], x[Np + m], x[2 * Np + m]}, t[3];
    sym_mul(b, xm, t);
    y[0] += t[0];
    y[1] += t[1];
    y[2] += t[2];
  });
}

template <bool FINE>
__global__ __launch_bounds__(256) void k_residual(LvArgs a, const float* x, const float* b, float* r, const int* gate) {
  if (gate && *gate) return;
  int n = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= a.N) return;
  const int Np = a.Np;
  if (a.fixed[n]) {
    r[n] = r[Np + n] = r[2 * Np + n] = 0.f;
    return;
  }
  float y[3];
  apply_row<FINE>(a, x, n, y);
  r[n] = b[n] - y[0];
  r[Np + n] = b[Np + n] - y[1];
  r[2 * Np + n] = b[2 * Np + n] - y[2];
}

// ---------------------------------------------------------------- prolongation (lattice.ts:13-41)
__device__ __forceinline__ int ccount(int n) { return n / 2 + 1; }  // ceil((n-1)/2) + 1
__device__ __forceinline__ int c2f(int I, int n) { return 2 * I < n - 1 ? 2 * I : n - 1; }

// bracket of fine index i (extent n): first coarse index, count (1 or 2), weights
__device__ __forceinline__ int bracket(int i, int n, int& lo, double w[2]) {
  int nc = ccount(n);
  lo = (i >> 1) < nc - 1 ? (i >> 1) : nc - 1;
  if (c2f(lo, n) > i) lo--;
  int a = c2f(lo, n);
  if (a == i) {
    w[0] = 1.0;
    return 1;
  }
  int b = c2f(lo + 1, n);
  w[0] = (double)(b - i) / (b - a);
  w[1] = (double)(i - a) / (b - a);
  return 2;
}

struct PwArgs {
  LvArgs a;
  Geo cgeo;
  float* pw;
  bool fine;
  const uint16_t* flags;
  const double* intact;
};

__device__ bool linked(const PwArgs& P, int a, int i, int j, int nx, int ny, int di, int dj) {
  int ii = i + di, jj = j + dj;
  if (ii < 0 || jj < 0 || ii >= nx || jj >= ny) return false;
  int b = a + dj * nx + di;
  if (P.fine) {
    int s = di > 0 ? 2 * a : di < 0 ? 2 * b : dj > 0 ? 2 * a + 1 : 2 * b + 1;
    return (P.flags[s] & (SEG_EXISTS | SEG_BROKEN)) == SEG_EXISTS && P.intact[s] > kLinkMinIntact;
  }
  if (P.a.fixed[b]) return true;
  int k = coarse_block(di, dj);
  int row = a;
  if (k < 0) {
    k = coarse_block(-di, -dj);
    row = b;
  }
  const int Np = P.a.Np;
  for (int q = 0; q < 6; q++)
    if (P.a.op[(6 * k + q) * Np + row] != 0.f) return true;
  return false;
}

// buildProlongation (newtonMultigrid.ts:305-342)
__global__ __launch_bounds__(256) void k_prolong_weights(PwArgs P) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= P.a.N) return;
  const int Np = P.a.Np;
  double w[4] = {0, 0, 0, 0};
  if (!P.a.fixed[f]) {
    int l, i, j;
    node_ij(P.a.geo, f, l, i, j);
    int nx = P.a.geo.nx[l], ny = P.a.geo.ny[l];
    int loj, loi;
    double wj[2], wi[2];
    int nj = bracket(j, ny, loj, wj);
    int ni = bracket(i, nx, loi, wi);
    double sum = 0;
    int k = 0;
    for (int a = 0; a < nj; a++) {
      int dj = c2f(loj + a, ny) - j;
      for (int b = 0; b < ni; b++) {
        int di = c2f(loi + b, nx) - i;
        bool ok = true;
        if (di != 0 && dj != 0) {
          int m1 = f + di, m2 = f + dj * nx;
          ok = (linked(P, f, i, j, nx, ny, di, 0) && linked(P, m1, i + di, j, nx, ny, 0, dj)) ||
               (linked(P, f, i, j, nx, ny, 0, dj) && linked(P, m2, i, j + dj, nx, ny, di, 0));
        } else if (di != 0 || dj != 0) {
          ok = linked(P, f, i, j, nx, ny, di, dj);
        }
        w[k] = ok ? wi[b] * wj[a] : 0.0;
        sum += w[k];
        k++;
      }
    }
    for (int q = 0; q < k; q++) w[q] = sum > 0 ? w[q] / sum : 0.0;
  }
  for (int q = 0; q < 4; q++) P.pw[q * Np + f] = (float)w[q];
}

// ---------------------------------------------------------------- Galerkin P^T A P
struct RapArgs {
  LvArgs a;     // fine level k
  Geo cgeo;     // coarse level k+1
  int Nc, Npc;
  const float* pw;
  float* cop;
  uint8_t* cfixed;
};

// weight of coarse (I, J) in fine node a's bracket (a at fine (i, j) of extent nx, ny), or 0
__device__ __forceinline__ float pw_of(const float* pw, int Np, int a, int i, int j, int nx, int ny, int I, int J) {
  int loj, loi;
  double wj[2], wi[2];
  int nj = bracket(j, ny, loj, wj);
  int ni = bracket(i, nx, loi, wi);
  int ej = J - loj, ei = I - loi;
  if (ej < 0 || ej >= nj || ei < 0 || ei >= ni) return 0.f;
  return pw[(ej * ni + ei) * Np + a];
}

template <bool FINE>
__global__ __launch_bounds__(64) void k_rap(RapArgs R) {
  __shared__ float acc[kCoarseComps * 64];
  const int t = threadIdx.x;
  const int I = blockIdx.x * 64 + t;
  for (int q = 0; q < kCoarseComps; q++) acc[q * 64 + t] = 0.f;
  if (I >= R.Nc) return;
  int l, Ii, Ij;
  node_ij(R.cgeo, I, l, Ii, Ij);
  const int nx = R.a.geo.nx[l], ny = R.a.geo.ny[l], foff = R.a.geo.off[l];
  const int Np = R.a.Np;
  const int cfi = c2f(Ii, nx), cfj = c2f(Ij, ny);
  for (int aj = cfj - 1; aj <= cfj + 1; aj++) {
    if (aj < 0 || aj >= ny) continue;
    for (int ai = cfi - 1; ai <= cfi + 1; ai++) {
      if (ai < 0 || ai >= nx) continue;
      const int a = foff + aj * nx + ai;
      if (R.a.fixed[a]) continue;
      const float wa = pw_of(R.pw, Np, a, ai, aj, nx, ny, Ii, Ij);
      if (wa == 0.f) continue;
      auto add = [&](int di, int dj, int b, const float* blk) {
        if (R.a.fixed[b]) return;
        int bi = ai + di, bj = aj + dj;
        int loj, loi;
        double wj[2], wi[2];
        int nj = bracket(bj, ny, loj, wj);
        int ni = bracket(bi, nx, loi, wi);
        for (int ej = 0; ej < nj; ej++)
          for (int ei = 0; ei < ni; ei++) {
            float wb = R.pw[(ej * ni + ei) * Np + b];
            if (wb == 0.f) continue;
            int kb = coarse_block(loi + ei - Ii, loj + ej - Ij);
            if (kb < 0) continue;
            float w = wa * wb;
            for (int q = 0; q < 6; q++) acc[(6 * kb + q) * 64 + t] += w * blk[q];
          }
      };
      float c[6];
      blk_load(R.a.op, Np, 0, a, c);
      add(0, 0, a, c);
      for_couplings<FINE>(R.a.op, Np, a, ai, aj, nx, ny, add);
    }
  }
  const int Npc = R.Npc;
  // empty diagonal: identity row, fixed (newtonMultigrid.ts:378-385)
  bool fx = acc[0 * 64 + t] + acc[3 * 64 + t] + acc[5 * 64 + t] <= 0.f;
  R.cfixed[I] = fx ? 1 : 0;
  for (int q = 0; q < kCoarseComps; q++) {
    float v = acc[q * 64 + t];
    if (fx) v = (q == 0 || q == 3 || q == 5) ? 1.f : 0.f;
    R.cop[q * Npc + I] = v;
  }
}

// ---------------------------------------------------------------- restriction / prolongation
struct XferArgs {
  LvArgs a;  // fine
  Geo cgeo;
  int Nc, Npc;
  const float* pw;
  const uint8_t* cfixed;
  const float* r;  // fine residual
  float* bc;       // coarse rhs
  float* xc;       // coarse solution
  float* xf;       // fine solution
  const int* gate;
};

__global__ __launch_bounds__(256) void k_restrict(XferArgs X) {
  if (X.gate && *X.gate) return;
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= X.Nc) return;
  int l, Ii, Ij;
  node_ij(X.cgeo, I, l, Ii, Ij);
  const int nx = X.a.geo.nx[l], ny = X.a.geo.ny[l], foff = X.a.geo.off[l];
  const int Np = X.a.Np, Npc = X.Npc;
  float s[3] = {0.f, 0.f, 0.f};
  if (!X.cfixed[I]) {
    const int cfi = c2f(Ii, nx), cfj = c2f(Ij, ny);
    for (int aj = cfj - 1; aj <= cfj + 1; aj++) {
      if (aj < 0 || aj >= ny) continue;
      for (int ai = cfi - 1; ai <= cfi + 1; ai++) {
        if (ai < 0 || ai >= nx) continue;
        int a = foff + aj * nx + ai;
        if (X.a.fixed[a]) continue;
        float w = pw_of(X.pw, Np, a, ai, aj, nx, ny, Ii, Ij);
        if (w == 0.f) continue;
        s[0] += w * X.r[a];
        s[1] += w * X.r[Np + a];
        s[2] += w * X.r[2 * Np + a];
      }
    }
  }
  for (int q = 0; q < 3; q++) {
    X.bc[q * Npc + I] = s[q];
    X.xc[q * Npc + I] = 0.f;
  }
}

__global__ __launch_bounds__(256) void k_prolong_add(XferArgs X) {
  if (X.gate && *X.gate) return;
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= X.a.N) return;
  if (X.a.fixed[f]) return;
  int l, i, j;
  node_ij(X.a.geo, f, l, i, j);
  const int nx = X.a.geo.nx[l], ny = X.a.geo.ny[l];
  const int cnx = X.cgeo.nx[l], coff = X.cgeo.off[l];
  const int Np = X.a.Np, Npc = X.Npc;
  int loj, loi;
  double wj[2], wi[2];
  int nj = bracket(j, ny, loj, wj);
  int ni = bracket(i, nx, loi, wi);
  float s[3] = {0.f, 0.f, 0.f};
  for (int ej = 0; ej < nj; ej++)
    for (int ei = 0; ei < ni; ei++) {
      float w = X.pw[(ej * ni + ei) * Np + f];
      int c = coff + (loj + ej) * cnx + loi + ei;
      s[0] += w * X.xc[c];
      s[1] += w * X.xc[Npc + c];
      s[2] += w * X.xc[2 * Npc + c];
    }
  for (int q = 0; q < 3; q++) X.xf[q * Np + f] += s[q];
}

// ---------------------------------------------------------------- line smoother
struct M3 {
  float m[9];
};

__device__ __forceinline__ void sym_to_full(const float* s, float* o) {
  o[0] = s[0]; o[1] = s[1]; o[2] = s[2];
  o[3] = s[1]; o[4] = s[3]; o[5] = s[4];
  o[6] = s[2]; o[7] = s[4]; o[8] = s[5];
}

// inverse of the symmetric block shifted by PIVOT_SHIFT * max|diag| (newtonMultigrid.ts:414-423)
__device__ __forceinline__ void pivot_inv(const float* s, float* o) {
  float e = (float)kPivotShift * fmaxf(fabsf(s[0]), fmaxf(fabsf(s[3]), fabsf(s[5])));
  float a00 = s[0] + e, a01 = s[1], a02 = s[2], a11 = s[3] + e, a12 = s[4], a22 = s[5] + e;
  float c00 = a11 * a22 - a12 * a12, c01 = a12 * a02 - a01 * a22, c02 = a01 * a12 - a11 * a02;
  float det = a00 * c00 + a01 * c01 + a02 * c02;
  float id = 1.f / det;
  o[0] = c00 * id;
  o[1] = c01 * id;
  o[2] = c02 * id;
  o[3] = (a00 * a22 - a02 * a02) * id;
  o[4] = (a02 * a01 - a00 * a12) * id;
  o[5] = (a00 * a11 - a01 * a01) * id;
}

__device__ __forceinline__ void mm(const float* a, const float* b, float* o) {
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) o[r * 3 + c] = a[r * 3] * b[c] + a[r * 3 + 1] * b[3 + c] + a[r * 3 + 2] * b[6 + c];
}
// a^T b
__device__ __forceinline__ void mtm(const float* a, const float* b, float* o) {
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) o[r * 3 + c] = a[r] * b[c] + a[3 + r] * b[3 + c] + a[6 + r] * b[6 + c];
}
// a b^T
__device__ __forceinline__ void mmt(const float* a, const float* b, float* o) {
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) o[r * 3 + c] = a[r * 3] * b[c * 3] + a[r * 3 + 1] * b[c * 3 + 1] + a[r * 3 + 2] * b[c * 3 + 2];
}
__device__ __forceinline__ void mv(const float* a, const float* v, float* o) {
  o[0] = a[0] * v[0] + a[1] * v[1] + a[2] * v[2];
  o[1] = a[3] * v[0] + a[4] * v[1] + a[5] * v[2];
  o[2] = a[6] * v[0] + a[7] * v[1] + a[8] * v[2];
}
__device__ __forceinline__ void mtv(const float* a, const float* v, float* o) {
  o[0] = a[0] * v[0] + a[3] * v[1] + a[6] * v[2];
  o[1] = a[1] * v[0] + a[4] * v[1] + a[7] * v[2];
  o[2] = a[2] * v[0] + a[5] * v[1] + a[8] * v[2];
}

struct SmoothArgs {
  LvArgs a;
  const int4* lines;
  float* x;
  const float* b;
  const int* gate;
};

// one colour of lines: block = one line, thread k = row k (smoothLines, newtonMultigrid.ts:659-733)
template <bool FINE>
__global__ __launch_bounds__(1024) void k_smooth(SmoothArgs S) {
  if (S.gate && *S.gate) return;
  extern __shared__ float sm[];
  const int4 ln = S.lines[blockIdx.x];
  const int len = ln.z, fam = ln.w;
  const int Lr = blockDim.x;
  float* sD = sm;           // 6 planes (sym D, later its shifted inverse)
  float* sU = sD + 6 * Lr;  // 9 planes: coupling to the next kept row
  float* sL = sU + 9 * Lr;  // 9 planes: coupling to the previous kept row (eliminated rows)
  float* sR = sL + 9 * Lr;  // 3 planes: rhs, then solution
  const int k = threadIdx.x;
  const int Np = S.a.Np;
  const bool row = k < len;
  int n = row ? ln.x + k * ln.y : 0;
  bool fixed_n = row ? S.a.fixed[n] != 0 : true;
  if (row) {
    float D[6], U[9], r[3];
    r[0] = S.b[n];
    r[1] = S.b[Np + n];
    r[2] = S.b[2 * Np + n];
    for (int q = 0; q < 9; q++) U[q] = 0.f;
    if (fixed_n) {
      D[0] = 1.f; D[1] = 0.f; D[2] = 0.f; D[3] = 1.f; D[4] = 0.f; D[5] = 1.f;
    } else {
      int l, i, j;
      node_ij(S.a.geo, n, l, i, j);
      blk_load(S.a.op, Np, 0, n, D);
      float maj = 0.f;
      for_couplings<FINE>(S.a.op, Np, n, i, j, S.a.geo.nx[l], S.a.geo.ny[l], [&](int di, int dj, int m, const float* blk) {
        if (S.a.fixed[m]) return;
        const int along = fam == 0 ? di : dj;
        const int across = fam == 0 ? dj : di;
        if (across == 0 && along == 1) {
          sym_to_full(blk, U);
          return;
        }
        if (across == 0 && along == -1) return;
        if (across == 0) maj += sym_frob(blk);  // +-2 along the line, majorised on the diagonal
        float xm[3] = {S.x[m], S.x[Np + m], S.x[2 * Np + m]}, t[3];
        sym_mul(blk, xm, t);
        r[0] -= t[0];
        r[1] -= t[1];
        r[2] -= t[2];
      });
      D[0] += maj;
      D[3] += maj;
      D[5] += maj;
      r[0] += maj * S.x[n];
      r[1] += maj * S.x[Np + n];
      r[2] += maj * S.x[2 * Np + n];
    }
    for (int q = 0; q < 6; q++) sD[q * Lr + k] = D[q];
    for (int q = 0; q < 9; q++) sU[q * Lr + k] = U[q];
    for (int q = 0; q < 3; q++) sR[q * Lr + k] = r[q];
  }
  __syncthreads();
  // odd-even block cyclic reduction
  int top = 1;
  while (2 * top < len) top <<= 1;
  for (int d = 1; d < len; d <<= 1) {
    const int d2 = 2 * d;
    if (row && (k & (d2 - 1)) == d) {  // eliminated at this level
      float s[6], inv[6];
      for (int q = 0; q < 6; q++) s[q] = sD[q * Lr + k];
      pivot_inv(s, inv);
      for (int q = 0; q < 6; q++) sD[q * Lr + k] = inv[q];
      // coupling to the previous kept row: transpose of its forward block
      float Up[9];
      for (int q = 0; q < 9; q++) Up[q] = sU[q * Lr + k - d];
      sL[0 * Lr + k] = Up[0]; sL[1 * Lr + k] = Up[3]; sL[2 * Lr + k] = Up[6];
      sL[3 * Lr + k] = Up[1]; sL[4 * Lr + k] = Up[4]; sL[5 * Lr + k] = Up[7];
      sL[6 * Lr + k] = Up[2]; sL[7 * Lr + k] = Up[5]; sL[8 * Lr + k] = Up[8];
    }
    __syncthreads();
    if (row && (k & (d2 - 1)) == 0) {  // kept
      float D[9], R[3], Unew[9];
      float Ds[6];
      for (int q = 0; q < 6; q++) Ds[q] = sD[q * Lr + k];
      sym_to_full(Ds, D);
      for (int q = 0; q < 3; q++) R[q] = sR[q * Lr + k];
      for (int q = 0; q < 9; q++) Unew[q] = 0.f;
      const int j1 = k - d, j2 = k + d;
      if (j1 >= 0) {
        float Di[6], Dif[9], U1[9], T[9], W[9], r1[3], y1[3], t[3];
        for (int q = 0; q < 6; q++) Di[q] = sD[q * Lr + j1];
        sym_to_full(Di, Dif);
        for (int q = 0; q < 9; q++) U1[q] = sU[q * Lr + j1];
        for (int q = 0; q < 3; q++) r1[q] = sR[q * Lr + j1];
        mm(Dif, U1, T);
        mtm(U1, T, W);
        for (int q = 0; q < 9; q++) D[q] -= W[q];
        mv(Dif, r1, y1);
        mtv(U1, y1, t);
        for (int q = 0; q < 3; q++) R[q] -= t[q];
      }
      if (j2 < len) {
        float Di[6], Dif[9], Uk[9], U2[9], T2[9], W[9], r2[3], t[3];
        for (int q = 0; q < 6; q++) Di[q] = sD[q * Lr + j2];
        sym_to_full(Di, Dif);
        for (int q = 0; q < 9; q++) Uk[q] = sU[q * Lr + k];
        for (int q = 0; q < 9; q++) U2[q] = sU[q * Lr + j2];
        for (int q = 0; q < 3; q++) r2[q] = sR[q * Lr + j2];
        mm(Uk, Dif, T2);
        mmt(T2, Uk, W);
        for (int q = 0; q < 9; q++) D[q] -= W[q];
        mv(T2, r2, t);
        for (int q = 0; q < 3; q++) R[q] -= t[q];
        mm(T2, U2, W);
        for (int q = 0; q < 9; q++) Unew[q] = -W[q];
      }
      // keep the symmetric part of the Schur complement
      sD[0 * Lr + k] = D[0];
      sD[1 * Lr + k] = 0.5f * (D[1] + D[3]);
      sD[2 * Lr + k] = 0.5f * (D[2] + D[6]);
      sD[3 * Lr + k] = D[4];
      sD[4 * Lr + k] = 0.5f * (D[5] + D[7]);
      sD[5 * Lr + k] = D[8];
      for (int q = 0; q < 3; q++) sR[q * Lr + k] = R[q];
      for (int q = 0; q < 9; q++) sU[q * Lr + k] = Unew[q];
    }
    __syncthreads();
  }
  if (k == 0 && len > 0) {
    float s[6], inv[6], Dif[9], R[3], x0[3];
    for (int q = 0; q < 6; q++) s[q] = sD[q * Lr];
    pivot_inv(s, inv);
    sym_to_full(inv, Dif);
    for (int q = 0; q < 3; q++) R[q] = sR[q * Lr];
    mv(Dif, R, x0);
    for (int q = 0; q < 3; q++) sR[q * Lr] = x0[q];
  }
  __syncthreads();
  for (int d = top; d >= 1; d >>= 1) {
    if (row && (k & (2 * d - 1)) == d) {
      float r[3], xl[3], t[3], Lk[9], Uk[9], Dif[9], Di[6], z[3];
      for (int q = 0; q < 3; q++) r[q] = sR[q * Lr + k];
      for (int q = 0; q < 3; q++) xl[q] = sR[q * Lr + k - d];
      for (int q = 0; q < 9; q++) Lk[q] = sL[q * Lr + k];
      mv(Lk, xl, t);
      for (int q = 0; q < 3; q++) r[q] -= t[q];
      if (k + d < len) {
        float xr[3];
        for (int q = 0; q < 3; q++) xr[q] = sR[q * Lr + k + d];
        for (int q = 0; q < 9; q++) Uk[q] = sU[q * Lr + k];
        mv(Uk, xr, t);
        for (int q = 0; q < 3; q++) r[q] -= t[q];
      }
      for (int q = 0; q < 6; q++) Di[q] = sD[q * Lr + k];
      sym_to_full(Di, Dif);
      mv(Dif, r, z);
      for (int q = 0; q < 3; q++) sR[q * Lr + k] = z[q];
    }
    __syncthreads();
  }
  if (row && !fixed_n) {
    S.x[n] = sR[k];
    S.x[Np + n] = sR[Lr + k];
    S.x[2 * Np + n] = sR[2 * Lr + k];
  }
}

__global__ void k_zero(float* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

LvArgs lv_args(const Level& L) { return LvArgs{L.op, L.fixed, L.N, L.Np, L.geo}; }

const int* pcg_gate(Solver& s) { return &s.scal->done; }

}  // namespace

void launch_zero(Solver& s, float* p, size_t n) {
  if (n == 0) return;
  k_zero<<<(unsigned)((n + 255) / 256), 256, 0, s.stream>>>(p, n);
}

void launch_prolong_weights(Solver& s, int k) {
  Level& L = s.lv[k];
  PwArgs P{lv_args(L), s.lv[k + 1].geo, L.pw, L.fine, s.flags, s.intact};
  ProfScope ps(s, PC_PROLONG_WEIGHTS, L.N * (L.fine ? 41.0 : 89.0));
  k_prolong_weights<<<grid_for(L.N, 256), 256, 0, s.stream>>>(P);
}

void launch_rap(Solver& s, int k) {
  Level& L = s.lv[k];
  Level& Cl = s.lv[k + 1];
  RapArgs R{lv_args(L), Cl.geo, Cl.N, Cl.Np, L.pw, Cl.op, Cl.fixed};
  ProfScope ps(s, PC_RAP, L.N * ((L.fine ? 232.0 : 600.0) + 17.0) + Cl.N * 313.0);
  if (L.fine)
    k_rap<true><<<grid_for(Cl.N, 64), 64, 0, s.stream>>>(R);
  else
    k_rap<false><<<grid_for(Cl.N, 64), 64, 0, s.stream>>>(R);
}

void launch_residual(Solver& s, int k, const float* x, const float* b, float* r) {
  Level& L = s.lv[k];
  ProfScope ps(s, L.fine ? PC_RESIDUAL_FINE : PC_RESIDUAL_COARSE, L.N * (L.fine ? 268.0 : 636.0));
  if (L.fine)
    k_residual<true><<<grid_for(L.N, 256), 256, 0, s.stream>>>(lv_args(L), x, b, r, pcg_gate(s));
  else
    k_residual<false><<<grid_for(L.N, 256), 256, 0, s.stream>>>(lv_args(L), x, b, r, pcg_gate(s));
}

void launch_restrict(Solver& s, int k) {
  Level& L = s.lv[k];
  Level& Cl = s.lv[k + 1];
  XferArgs X{lv_args(L), Cl.geo, Cl.N, Cl.Np, L.pw, Cl.fixed, L.r, Cl.b, Cl.x, L.x, pcg_gate(s)};
  ProfScope ps(s, PC_RESTRICT, L.N * 29.0 + Cl.N * 25.0);
  k_restrict<<<grid_for(Cl.N, 256), 256, 0, s.stream>>>(X);
}

void launch_prolong_add(Solver& s, int k) {
  Level& L = s.lv[k];
  Level& Cl = s.lv[k + 1];
  XferArgs X{lv_args(L), Cl.geo, Cl.N, Cl.Np, L.pw, Cl.fixed, L.r, Cl.b, Cl.x, L.x, pcg_gate(s)};
  ProfScope ps(s, PC_PROLONG, L.N * 41.0 + Cl.N * 12.0);
  k_prolong_add<<<grid_for(L.N, 256), 256, 0, s.stream>>>(X);
}

void launch_smooth(Solver& s, int k, int reverse) {
  Level& L = s.lv[k];
  for (int q = 0; q < 6; q++) {
    int c = reverse ? 5 - q : q;
    int count = L.color_start[c + 1] - L.color_start[c];
    if (count == 0) continue;
    int maxlen = 0;
    // lines of a colour share one family, so one length bound per colour
    maxlen = c < 3 ? 0 : 0;
    for (int l = 0; l < L.geo.nl; l++) maxlen = std::max(maxlen, c < 3 ? L.geo.ny[l] : L.geo.nx[l]);
    int block = ((maxlen + 63) / 64) * 64;
    size_t shm = (size_t)27 * block * sizeof(float);
    SmoothArgs S{lv_args(L), L.lines + L.color_start[c], L.x, L.b, pcg_gate(s)};
    // each node of the colour's lines: its full row (fine 58 floats, coarse 150) + b, x read, x write
    double nodes = (double)count * (c < 3 ? L.geo.ny[0] : L.geo.nx[0]);
    ProfScope ps(s, L.fine ? PC_SMOOTH_FINE : PC_SMOOTH_COARSE, nodes * (L.fine ? 268.0 : 636.0));
    if (L.fine)
      k_smooth<true><<<count, block, shm, s.stream>>>(S);
    else
      k_smooth<false><<<count, block, shm, s.stream>>>(S);
  }
}

void smooth_set_attributes() {
  int shm = 27 * BD_MAX_LINE * (int)sizeof(float);
  BD_HIP(hipFuncSetAttribute((const void*)k_smooth<true>, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
  BD_HIP(hipFuncSetAttribute((const void*)k_smooth<false>, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
}

// vcycle (newtonMultigrid.ts:495-528) on level k's x / b buffers, x zero on entry
void vcycle(Solver& s, int k) {
  Level& L = s.lv[k];
  if (k == s.nlev - 1) {
    for (int sw = 0; sw < s.P.coarse_sweeps; sw++) launch_smooth(s, k, sw & 1);
    return;
  }
  launch_smooth(s, k, 0);
  launch_residual(s, k, L.x, L.b, L.r);
  launch_restrict(s, k);
  vcycle(s, k + 1);
  launch_prolong_add(s, k);
  launch_smooth(s, k, 1);
}

}  // namespace bd

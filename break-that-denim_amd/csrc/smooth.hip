// Colour-ordered block-tridiagonal yarn-line smoother (smoothLines / solvePenta,
// src/simulation/reference/newtonMultigrid.ts:435-473, 659-733; line table lattice.ts:86-106).
//
// One colour pass solves every line of that colour exactly: D_k x_k + U_k x_{k+1} + U_{k-1}^T x_{k-1}
// = r_k, where D_k is the node's centre block (plus the Frobenius norms of its +-2 along-line
// couplings, majorised onto the diagonal), U_k its +1 along-line coupling, and r_k its rhs minus
// every other coupling times the lagged iterate. Lines of one colour share no coupling (3 colours),
// so reading x in place equals the reference's per-colour snapshot.
//
// MI355X mapping: one workgroup per line, two rows per thread. The row gather is branch-free (all
// ~110 loads of a row issue at once, clamped indices with masks) because a colour pass has only a
// few hundred lines in flight and is latency-bound. The first odd-even reduction level runs in
// registers (the thread owns both rows of a pair; neighbours exchange one 9-float term via LDS), the
// remaining log2(len) - 1 levels by block parallel cyclic reduction, all reduced rows active at every
// level (one barrier per level and no back-substitution: half the serial steps of odd-even cyclic
// reduction, which is what bounds a pass). LDS: two 20-float records per reduced row (80 KB for a
// 1024-node line, two workgroups per CU). Every pivot is shifted by PIVOT_SHIFT of its largest diagonal entry,
// as in the reference.
//
// Levels with lines of kDualMinLine (33) or more nodes also keep their operator, fixed flags and iterate in a
// column-major copy (node (i, j) of a layer at off + i * ny + j): v-line passes read that copy, so
// consecutive rows of a v-line are consecutive in memory and the gathers coalesce as they do for
// u-lines. Each pass writes its solution to both copies (the other one with a 12-byte strided
// store per row) and reads b from the row-major vector.
#include "bd_internal.h"
#include "oprd.h"

namespace bd {

namespace {

struct SmoothArgs {
  const float* op;       // in the pass's layout (column-major for v-lines of a dual-layout level)
  const uint8_t* fixed;  // same layout
  int Np;
  const int4* lines;     // (start, stride, length, family), start in the pass's layout
  float* x;              // same layout
  const float* b;        // same layout
  float* xo;             // the other layout's copy of x (dual-layout levels), or null
  const int* gate;
  Geo geo;
};

// index strides of lattice offsets (di, dj) in the pass's layout
template <int FAM, bool DUAL>
__device__ __forceinline__ int stride_i(int nx, int ny) { return (FAM == 1 && DUAL) ? ny : 1; }
template <int FAM, bool DUAL>
__device__ __forceinline__ int stride_j(int nx, int ny) { return (FAM == 1 && DUAL) ? 1 : nx; }

__device__ __forceinline__ void sym_full(const float* s, float* o) {
  o[0] = s[0]; o[1] = s[1]; o[2] = s[2];
  o[3] = s[1]; o[4] = s[3]; o[5] = s[4];
  o[6] = s[2]; o[7] = s[4]; o[8] = s[5];
}

// inverse of the symmetric block shifted by PIVOT_SHIFT * max|diag| (newtonMultigrid.ts:414-423)
__device__ __forceinline__ void pivot_inv(const float* s, float* o) {
  float e = (float)kPivotShift * fmaxf(fabsf(s[0]), fmaxf(fabsf(s[3]), fabsf(s[5])));
  float a00 = s[0] + e, a01 = s[1], a02 = s[2], a11 = s[3] + e, a12 = s[4], a22 = s[5] + e;
  float c00 = a11 * a22 - a12 * a12, c01 = a12 * a02 - a01 * a22, c02 = a01 * a12 - a11 * a02;
  // hardware reciprocal (1 ulp): the pivots of a preconditioner need no correctly rounded division
  float id = __builtin_amdgcn_rcpf(a00 * c00 + a01 * c01 + a02 * c02);
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
__device__ __forceinline__ void mtm(const float* a, const float* b, float* o) {  // a^T b
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) o[r * 3 + c] = a[r] * b[c] + a[3 + r] * b[3 + c] + a[6 + r] * b[6 + c];
}
__device__ __forceinline__ void mmt(const float* a, const float* b, float* o) {  // a b^T
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
__device__ __forceinline__ void smv(const float* s, const float* v, float* o) {
  o[0] = s[0] * v[0] + s[1] * v[1] + s[2] * v[2];
  o[1] = s[1] * v[0] + s[3] * v[1] + s[4] * v[2];
  o[2] = s[2] * v[0] + s[4] * v[1] + s[5] * v[2];
}

// one coupling of the row gather: source row and plane base of its symmetric block (or scalar)
struct Cpl {
  int di, dj;   // offset of the neighbour
  int sdi, sdj; // offset of the row that stores the block (0,0: own forward planes)
  int base;     // first plane
  bool scalar;
};

// fine stencil: 4 forward blocks, their 4 mirrors, 2 forward scalars and their mirrors
__device__ constexpr Cpl kFine[12] = {
    {1, 0, 0, 0, F_E, false},   {-1, 0, -1, 0, F_E, false}, {0, 1, 0, 0, F_N, false},    {0, -1, 0, -1, F_N, false},
    {1, 1, 0, 0, F_NE, false},  {-1, -1, -1, -1, F_NE, false}, {-1, 1, 0, 0, F_NW, false}, {1, -1, 1, -1, F_NW, false},
    {2, 0, 0, 0, F_EE, true},   {-2, 0, -2, 0, F_EE, true}, {0, 2, 0, 0, F_NN, true},    {0, -2, 0, -2, F_NN, true}};

// Gathers D (6), U (9: +1 along coupling, full), r (3) of row n on a line of family FAM.
// All loads use clamped in-bounds indices and are masked afterwards, so they issue together.
template <bool FINE, int FAM, bool DUAL, class RD>
__device__ __forceinline__ void gather_row(const SmoothArgs& S, const RD& rd, int n, int nb, int i, int j, int nx, int ny,
                                           float D[6], float U[9], float r[3], bool& fixed_n) {
  const int Np = S.Np;
  const float* x = S.x;
  const int si = stride_i<FAM, DUAL>(nx, ny), sj = stride_j<FAM, DUAL>(nx, ny);
  fixed_n = S.fixed[n] != 0;
  float c[6];
  rd.blk(F_C, n, c);
  const float xn0 = x[n], xn1 = x[Np + n], xn2 = x[2 * Np + n];
  r[0] = S.b[n];
  r[1] = S.b[Np + n];
  r[2] = S.b[2 * Np + n];
#pragma unroll
  for (int q = 0; q < 9; q++) U[q] = 0.f;
  float maj = 0.f;
  auto one = [&](int di, int dj, int sdi, int sdj, int base, bool scalar, bool neg_src) {
    const bool in = i + di >= 0 && i + di < nx && j + dj >= 0 && j + dj < ny;
    const int m = in ? n + dj * sj + di * si : n;
    const int src = in ? n + sdj * sj + sdi * si : n;
    float blk[6];
    if (scalar) {
      float s = rd.scalar(base, src);
      blk[0] = s; blk[1] = 0.f; blk[2] = 0.f; blk[3] = s; blk[4] = 0.f; blk[5] = s;
    } else {
      rd.blk(base, src, blk);
    }
    const float xm[3] = {x[m], x[Np + m], x[2 * Np + m]};
    // the fixed-neighbour mask is arithmetic: a branch on the loaded flag would split the gather
    // into one memory round trip per coupling
    const uint8_t fm = S.fixed[m];  // m is clamped in bounds: the load is unconditional
    const float w = (in & (fm == 0)) ? 1.f : 0.f;
    const int along = FAM == 0 ? di : dj;
    const int across = FAM == 0 ? dj : di;
    if (across == 0 && along == 1) {
      float full[9];
      sym_full(blk, full);
#pragma unroll
      for (int q = 0; q < 9; q++) U[q] = w * full[q];
    } else if (!(across == 0 && along == -1)) {
      float t[3];
      smv(blk, xm, t);
      r[0] -= w * t[0];
      r[1] -= w * t[1];
      r[2] -= w * t[2];
      if (across == 0) maj += w * sym_frob(blk);  // +-2 along the line
    }
  };
  if (FINE) {
#pragma unroll
    for (int k = 0; k < 12; k++) one(kFine[k].di, kFine[k].dj, kFine[k].sdi, kFine[k].sdj, kFine[k].base, kFine[k].scalar, false);
  } else {
    // 24 couplings of 6-plane blocks: batches of 6 keep the loads in flight within the VGPR budget
#pragma unroll 3
    for (int k = 1; k < kCoarseBlocks; k++) {
      one(coarse_di(k), coarse_dj(k), 0, 0, 6 * k, false, false);
      one(-coarse_di(k), -coarse_dj(k), -coarse_di(k), -coarse_dj(k), 6 * k, false, true);
    }
  }
  if (fixed_n) {
    D[0] = 1.f; D[1] = 0.f; D[2] = 0.f; D[3] = 1.f; D[4] = 0.f; D[5] = 1.f;
#pragma unroll
    for (int q = 0; q < 9; q++) U[q] = 0.f;
    r[0] = S.b[n];
    r[1] = S.b[Np + n];
    r[2] = S.b[2 * Np + n];
    return;
  }
#pragma unroll
  for (int q = 0; q < 6; q++) D[q] = c[q];
  D[0] += maj;
  D[3] += maj;
  D[5] += maj;
  r[0] += maj * xn0;
  r[1] += maj * xn1;
  r[2] += maj * xn2;
}

// Partial row gather of rows on one wave's 64 lanes for the couplings c with c % G == GI (GI is a
// compile-time constant, so the part is straight-line code whose loads all issue together). GI 0
// also loads the row's centre block, b, x and fixed flag. Writes the partial rhs correction and
// majorisation to LDS; the part owning the +1 along-line coupling writes U. Lanes past the line's
// end (valid false) gather a clamped row and write nothing.
template <bool FINE, int FAM, int G, int GI, bool DUAL, class RD>
__device__ __forceinline__ void gather_part(const SmoothArgs& S, const RD& rd, int n, int nb, int i, int j, int nx, int ny,
                                            bool valid, int row, int len, float* pr, float* pm, float* gU, float* gD,
                                            float* gR, float* gX, uint8_t* gF) {
  const int Np = S.Np;
  const float* x = S.x;
  const int si = stride_i<FAM, DUAL>(nx, ny), sj = stride_j<FAM, DUAL>(nx, ny);
  float r0 = 0.f, r1 = 0.f, r2 = 0.f, maj = 0.f;
  float U[9];
  bool ownsU = false;
  auto one = [&](int di, int dj, int sdi, int sdj, int base, bool scalar) {
    const bool in = i + di >= 0 && i + di < nx && j + dj >= 0 && j + dj < ny;
    const int m = in ? n + dj * sj + di * si : n;
    const int src = in ? n + sdj * sj + sdi * si : n;
    float blk[6];
    if (scalar) {
      float s = rd.scalar(base, src);
      blk[0] = s; blk[1] = 0.f; blk[2] = 0.f; blk[3] = s; blk[4] = 0.f; blk[5] = s;
    } else {
      rd.blk(base, src, blk);
    }
    const float xm[3] = {x[m], x[Np + m], x[2 * Np + m]};
    const uint8_t fm = S.fixed[m];  // m is clamped in bounds: the load is unconditional
    const float w = (in & (fm == 0)) ? 1.f : 0.f;
    const int along = FAM == 0 ? di : dj;
    const int across = FAM == 0 ? dj : di;
    if (across == 0 && along == 1) {
      float full[9];
      sym_full(blk, full);
#pragma unroll
      for (int q = 0; q < 9; q++) U[q] = w * full[q];
      ownsU = true;
    } else if (!(across == 0 && along == -1)) {
      float t[3];
      smv(blk, xm, t);
      r0 += w * t[0];
      r1 += w * t[1];
      r2 += w * t[2];
      if (across == 0) maj += w * sym_frob(blk);
    }
  };
  if (FINE) {
#pragma unroll
    for (int k = 0; k < 12; k++)
      if (k % G == GI) one(kFine[k].di, kFine[k].dj, kFine[k].sdi, kFine[k].sdj, kFine[k].base, kFine[k].scalar);
  } else {
#pragma unroll
    for (int k = 0; k < 24; k++) {
      if (k % G != GI) continue;
      const int b = 1 + (k >> 1);
      if (k & 1)
        one(-coarse_di(b), -coarse_dj(b), -coarse_di(b), -coarse_dj(b), 6 * b, false);
      else
        one(coarse_di(b), coarse_dj(b), 0, 0, 6 * b, false);
    }
  }
  float c[6], bb[3], xx[3];
  uint8_t f = 0;
  if (GI == 0) {
    rd.blk(F_C, n, c);
#pragma unroll
    for (int q = 0; q < 3; q++) {
      bb[q] = S.b[q * Np + n];
      xx[q] = x[q * Np + n];
    }
    f = S.fixed[n];
  }
  if (!valid) return;
  pr[(0 * G + GI) * len + row] = r0;
  pr[(1 * G + GI) * len + row] = r1;
  pr[(2 * G + GI) * len + row] = r2;
  pm[GI * len + row] = maj;
  if (ownsU) {
#pragma unroll
    for (int q = 0; q < 9; q++) gU[q * len + row] = U[q];
  }
  if (GI == 0) {
#pragma unroll
    for (int q = 0; q < 6; q++) gD[q * len + row] = c[q];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      gR[q * len + row] = bb[q];
      gX[q * len + row] = xx[q];
    }
    gF[row] = f;
  }
}

// Runs part g (wave-uniform) of the gather.
template <bool FINE, int FAM, int G, bool DUAL, class RD, int GI = 0>
__device__ __forceinline__ void gather_dispatch(int g, const SmoothArgs& S, const RD& rd, int n, int nb, int i, int j, int nx,
                                                int ny, bool valid, int row, int len, float* pr, float* pm, float* gU,
                                                float* gD, float* gR, float* gX, uint8_t* gF) {
  if constexpr (GI < G) {
    if (g == GI)
      gather_part<FINE, FAM, G, GI, DUAL>(S, rd, n, nb, i, j, nx, ny, valid, row, len, pr, pm, gU, gD, gR, gX, gF);
    else
      gather_dispatch<FINE, FAM, G, DUAL, RD, GI + 1>(g, S, rd, n, nb, i, j, nx, ny, valid, row, len, pr, pm, gU, gD, gR,
                                                      gX, gF);
  }
}

// LDS record of one reduced row in the line solve: D^-1 (6, symmetric), U (9), r (3) and 2 floats of
// padding, so that a record is five 16-byte vector accesses with immediate offsets
constexpr int kRowF = 20;
// LDS floats per reduced row: two record buffers
constexpr int kSolveFloats = 2 * kRowF;
// variant code of the split smoother (row kernel + solve kernel)
constexpr int kSplitVariant = 16;

__device__ __forceinline__ void put_row(float* B, int t, const float Di[6], const float U[9], const float r[3]) {
  float4* p = reinterpret_cast<float4*>(B + kRowF * t);
  p[0] = make_float4(Di[0], Di[1], Di[2], Di[3]);
  p[1] = make_float4(Di[4], Di[5], U[0], U[1]);
  p[2] = make_float4(U[2], U[3], U[4], U[5]);
  p[3] = make_float4(U[6], U[7], U[8], r[0]);
  p[4] = make_float4(r[1], r[2], 0.f, 0.f);
}

__device__ __forceinline__ void get_row(const float* B, int j, float Di[6], float U[9], float r[3]) {
  const float4* p = reinterpret_cast<const float4*>(B + kRowF * j);
  const float4 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4];
  Di[0] = a.x; Di[1] = a.y; Di[2] = a.z; Di[3] = a.w; Di[4] = b.x; Di[5] = b.y;
  U[0] = b.z; U[1] = b.w; U[2] = c.x; U[3] = c.y; U[4] = c.z; U[5] = c.w; U[6] = d.x; U[7] = d.y; U[8] = d.z;
  r[0] = d.w; r[1] = e.x; r[2] = e.y;
}

// Block tridiagonal solve of one line, two rows per thread.
// Thread t holds rows a = 2t and b = 2t+1 (D sym, U full coupling to the next row, r) and
// receives their solutions. The first odd-even level eliminates row b in registers (neighbours
// exchange one 9-float term through LDS), leaving R2 = ceil(len/2) reduced rows, one per thread.
// Those are solved by block parallel cyclic reduction: at level d every row eliminates both its
// neighbours at distance d at once, so after ceil(log2 R2) levels (one barrier each, the row kept
// in registers, LDS ping-ponged) each row stands alone; there is no back-substitution sweep. The
// reduced system stays symmetric (row k's coupling to k - d is row k - d's coupling to k,
// transposed), so a row carries only its forward coupling, and each row publishes its own shifted
// pivot inverse with it (one inversion per row and level). cr: kSolveFloats * Lr floats of LDS.
__device__ __forceinline__ void line_solve(float* cr, int Lr, int t, int len, const float* Da_in, const float* Ua,
                                           const float* ra_in, const float* Db, const float* Ub, const float* rb,
                                           float xa[3], float xb[3]) {
  float* B0 = cr;
  float* B1 = cr + kRowF * Lr;  // also carries the register level's exchange (its first 9 floats)
  const int R2 = (len + 1) >> 1;
  const bool has_a = 2 * t < len, has_b = 2 * t + 1 < len;
  float D[6], r[3], U[9];
#pragma unroll
  for (int q = 0; q < 6; q++) D[q] = Da_in[q];
#pragma unroll
  for (int q = 0; q < 3; q++) r[q] = ra_in[q];

  // register level: eliminate row b (U_a couples a -> b, U_b couples b -> 2t+2)
  float Dib[6], Dibf[9], yb[3];
  pivot_inv(Db, Dib);
  sym_full(Dib, Dibf);
  smv(Dib, rb, yb);
  {
    float T[9], P[9], p[3];
    mm(Dibf, Ub, T);
    mtm(Ub, T, P);
    mtv(Ub, yb, p);
    if (has_b && t + 1 < R2) {
      float4* e = reinterpret_cast<float4*>(B1 + kRowF * t);
      e[0] = make_float4(P[0], 0.5f * (P[1] + P[3]), 0.5f * (P[2] + P[6]), P[4]);
      e[1] = make_float4(0.5f * (P[5] + P[7]), P[8], p[0], p[1]);
      e[2] = make_float4(p[2], 0.f, 0.f, 0.f);
    }
  }
  {
    float T2[9], W[9], t3[3];
    mm(Ua, Dibf, T2);
    mmt(T2, Ua, W);
    D[0] -= W[0];
    D[1] -= 0.5f * (W[1] + W[3]);
    D[2] -= 0.5f * (W[2] + W[6]);
    D[3] -= W[4];
    D[4] -= 0.5f * (W[5] + W[7]);
    D[5] -= W[8];
    mv(T2, rb, t3);
    r[0] -= t3[0];
    r[1] -= t3[1];
    r[2] -= t3[2];
    mm(T2, Ub, W);
#pragma unroll
    for (int q = 0; q < 9; q++) U[q] = -W[q];
  }
  __syncthreads();
  if (has_a && t >= 1) {
    const float4* e = reinterpret_cast<const float4*>(B1 + kRowF * (t - 1));
    const float4 e0 = e[0], e1 = e[1], e2 = e[2];
    D[0] -= e0.x; D[1] -= e0.y; D[2] -= e0.z; D[3] -= e0.w; D[4] -= e1.x; D[5] -= e1.y;
    r[0] -= e1.z; r[1] -= e1.w; r[2] -= e2.x;
  }
  // B1's exchange terms are read before the first level's barrier; that level writes B0 and the
  // next one B1, so one barrier per level orders every read before the overwrite.
  const bool row = t < R2;
  int lvl = 0;
  for (int d = 1; d < R2; d <<= 1, lvl ^= 1) {
    float* B = lvl ? B1 : B0;
    if (row) {
      float Di[6];
      pivot_inv(D, Di);
      put_row(B, t, Di, U, r);
    }
    __syncthreads();
    if (row) {
      float Un[9];
#pragma unroll
      for (int q = 0; q < 9; q++) Un[q] = 0.f;
      const int j1 = t - d, j2 = t + d;
      if (j1 >= 0) {
        // row j1 couples to this row by U1: D -= U1^T D1^-1 U1, r -= U1^T D1^-1 r1
        float Di[6], U1[9], r1[3], T[9];
        get_row(B, j1, Di, U1, r1);
        float Dif[9];
        sym_full(Di, Dif);
        mm(Dif, U1, T);  // T = D1^-1 U1
        D[0] -= U1[0] * T[0] + U1[3] * T[3] + U1[6] * T[6];
        D[1] -= U1[0] * T[1] + U1[3] * T[4] + U1[6] * T[7];
        D[2] -= U1[0] * T[2] + U1[3] * T[5] + U1[6] * T[8];
        D[3] -= U1[1] * T[1] + U1[4] * T[4] + U1[7] * T[7];
        D[4] -= U1[1] * T[2] + U1[4] * T[5] + U1[7] * T[8];
        D[5] -= U1[2] * T[2] + U1[5] * T[5] + U1[8] * T[8];
        float tt[3];
        mtv(T, r1, tt);  // U1^T D1^-1 r1
        r[0] -= tt[0];
        r[1] -= tt[1];
        r[2] -= tt[2];
      }
      if (j2 < R2) {
        // this row couples to row j2 by U, row j2 on to j2 + d by U2: D -= U D2^-1 U^T,
        // r -= U D2^-1 r2, and the new coupling to j2 + d is -U D2^-1 U2
        float Di[6], U2[9], r2[3], T2[9];
        get_row(B, j2, Di, U2, r2);
        float Dif[9];
        sym_full(Di, Dif);
        mm(U, Dif, T2);  // T2 = U D2^-1
        D[0] -= T2[0] * U[0] + T2[1] * U[1] + T2[2] * U[2];
        D[1] -= T2[0] * U[3] + T2[1] * U[4] + T2[2] * U[5];
        D[2] -= T2[0] * U[6] + T2[1] * U[7] + T2[2] * U[8];
        D[3] -= T2[3] * U[3] + T2[4] * U[4] + T2[5] * U[5];
        D[4] -= T2[3] * U[6] + T2[4] * U[7] + T2[5] * U[8];
        D[5] -= T2[6] * U[6] + T2[7] * U[7] + T2[8] * U[8];
        float tt[3];
        mv(T2, r2, tt);
        r[0] -= tt[0];
        r[1] -= tt[1];
        r[2] -= tt[2];
        float W[9];
        mm(T2, U2, W);
#pragma unroll
        for (int q = 0; q < 9; q++) Un[q] = -W[q];
      }
#pragma unroll
      for (int q = 0; q < 9; q++) U[q] = Un[q];
    }
  }
  // every reduced row now stands alone
  float* B = lvl ? B1 : B0;
  xa[0] = xa[1] = xa[2] = 0.f;
  if (row) {
    float inv[6];
    pivot_inv(D, inv);
    smv(inv, r, xa);
    B[kRowF * t] = xa[0];
    B[kRowF * t + 1] = xa[1];
    B[kRowF * t + 2] = xa[2];
  }
  __syncthreads();
  float xn[3] = {0.f, 0.f, 0.f};
  if (t + 1 < R2) {
    xn[0] = B[kRowF * (t + 1)];
    xn[1] = B[kRowF * (t + 1) + 1];
    xn[2] = B[kRowF * (t + 1) + 2];
  }
  float tt[3], q1[3];
  mtv(Ua, xa, tt);
  mv(Ub, xn, q1);
  float rr[3] = {rb[0] - tt[0] - q1[0], rb[1] - tt[1] - q1[1], rb[2] - tt[2] - q1[2]};
  smv(Dib, rr, xb);
}

__device__ __forceinline__ void set_identity(float D[6], float U[9], float r[3]) {
  D[0] = 1.f; D[1] = 0.f; D[2] = 0.f; D[3] = 1.f; D[4] = 0.f; D[5] = 1.f;
#pragma unroll
  for (int q = 0; q < 9; q++) U[q] = 0.f;
  r[0] = r[1] = r[2] = 0.f;
}

struct LineGeo {
  int4 ln;
  int nx, ny, off, i0, j0;
};

template <int FAM, bool DUAL>
__device__ __forceinline__ LineGeo line_geo_at(const SmoothArgs& S, int idx) {
  LineGeo g;
  g.ln = S.lines[idx];
  int l = geo_layer(S.geo, g.ln.x);
  g.nx = S.geo.nx[l];
  g.ny = S.geo.ny[l];
  g.off = S.geo.off[l];
  int r0 = g.ln.x - g.off;
  if (FAM == 1 && DUAL) {
    g.i0 = r0 / g.ny;
    g.j0 = r0 - g.i0 * g.ny;
  } else {
    g.j0 = r0 / g.nx;
    g.i0 = r0 - g.j0 * g.nx;
  }
  return g;
}

// the line of this workgroup (blocks spread over the XCDs)
template <int FAM, bool DUAL>
__device__ __forceinline__ LineGeo line_geo(const SmoothArgs& S) {
  return line_geo_at<FAM, DUAL>(S, xcd_block(blockIdx.x, gridDim.x));
}

// row k of the line: its index in the pass's layout, in the row-major layout (b), and in the other
// layout (the second copy of x on dual-layout levels)
template <int FAM, bool DUAL>
__device__ __forceinline__ void row_index(const LineGeo& L, int k, int& n, int& nb, int& no, int& i, int& j) {
  n = L.ln.x + k * L.ln.y;
  i = FAM == 0 ? L.i0 + k : L.i0;
  j = FAM == 0 ? L.j0 : L.j0 + k;
  const int rm = L.off + j * L.nx + i, cm = L.off + i * L.ny + j;
  nb = (FAM == 1 && DUAL) ? rm : n;
  no = DUAL ? (FAM == 1 ? rm : cm) : n;
}

__device__ __forceinline__ void store_x(const SmoothArgs& S, int n, int no, const float v[3], bool dual) {
  const int Np = S.Np;
  S.x[n] = v[0];
  S.x[Np + n] = v[1];
  S.x[2 * Np + n] = v[2];
  if (dual) {
    S.xo[no] = v[0];
    S.xo[Np + no] = v[1];
    S.xo[2 * Np + no] = v[2];
  }
}

// Long lines: each thread gathers its two rows straight into registers.
template <bool FINE, int FAM, bool DUAL, class RDD>
__global__ __launch_bounds__(512, 4) void k_smooth_direct(SmoothArgs S, RDD rdd) {
  if (S.gate && *S.gate) return;
  const auto rd = rdd.live();
  extern __shared__ __align__(16) float sm[];
  const LineGeo L = line_geo<FAM, DUAL>(S);
  const int len = L.ln.z;
  const int t = threadIdx.x;
  const int ka = 2 * t, kb = 2 * t + 1;
  int na, nba, noa, ia, ja, nb, nbb, nob, ib, jb;
  row_index<FAM, DUAL>(L, ka, na, nba, noa, ia, ja);
  row_index<FAM, DUAL>(L, kb, nb, nbb, nob, ib, jb);
  float Da[6], Ua[9], ra[3], Db[6], Ub[9], rb[3];
  bool fa = true, fb = true;
  if (ka < len)
    gather_row<FINE, FAM, DUAL>(S, rd, na, nba, ia, ja, L.nx, L.ny, Da, Ua, ra, fa);
  else
    set_identity(Da, Ua, ra);
  if (kb < len)
    gather_row<FINE, FAM, DUAL>(S, rd, nb, nbb, ib, jb, L.nx, L.ny, Db, Ub, rb, fb);
  else
    set_identity(Db, Ub, rb);
  float xa[3], xb[3];
  line_solve(sm, blockDim.x, t, len, Da, Ua, ra, Db, Ub, rb, xa, xb);
  if (ka < len && !fa) store_x(S, na, noa, xa, DUAL);
  if (kb < len && !fb) store_x(S, nb, nob, xb, DUAL);
}

// Short lines: every row's couplings are split over G threads so that each thread issues about
// one round trip of loads; rows are assembled in LDS, then solved by the first ceil(len/2) threads.
template <bool FINE, int FAM, int G, bool DUAL, class RDD>
__global__ __launch_bounds__(1024) void k_smooth_staged(SmoothArgs S, RDD rdd) {
  if (S.gate && *S.gate) return;
  const auto rd = rdd.live();
  extern __shared__ __align__(16) float sm[];
  const LineGeo L = line_geo<FAM, DUAL>(S);
  const int len = L.ln.z;
  const int R2 = (len + 1) >> 1;
  const int Lr = ((R2 + 63) / 64) * 64;
  float* cr = sm;              // kSolveFloats * Lr: line-solve workspace
  float* gD = cr + kSolveFloats * Lr;  // 6 * len: centre blocks, then D
  float* gU = gD + 6 * len;    // 9 * len
  float* gR = gU + 9 * len;    // 3 * len: b, then r
  float* gX = gR + 3 * len;    // 3 * len: own x
  float* pr = gX + 3 * len;    // 3 * G * len
  float* pm = pr + 3 * G * len;  // G * len
  uint8_t* gF = reinterpret_cast<uint8_t*>(pm + G * len);
  const int T = blockDim.x;
  // work items: (64 consecutive rows, coupling part g), one wave each, so that g is wave-uniform
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int items = ((len + 63) >> 6) * G;
  for (int item = wave; item < items; item += T >> 6) {
    const int g = item % G, row0 = (item / G) * 64 + lane;
    const bool valid = row0 < len;
    const int row = valid ? row0 : len - 1;
    int n, nb, no, i, j;
    row_index<FAM, DUAL>(L, row, n, nb, no, i, j);
    gather_dispatch<FINE, FAM, G, DUAL>(g, S, rd, n, nb, i, j, L.nx, L.ny, valid, row, len, pr, pm, gU, gD, gR, gX, gF);
  }
  __syncthreads();
  for (int row = threadIdx.x; row < len; row += T) {
    if (gF[row]) {
      gD[0 * len + row] = 1.f; gD[1 * len + row] = 0.f; gD[2 * len + row] = 0.f;
      gD[3 * len + row] = 1.f; gD[4 * len + row] = 0.f; gD[5 * len + row] = 1.f;
#pragma unroll
      for (int q = 0; q < 9; q++) gU[q * len + row] = 0.f;
      continue;
    }
    float maj = 0.f, s0 = 0.f, s1 = 0.f, s2 = 0.f;
#pragma unroll
    for (int g = 0; g < G; g++) {
      maj += pm[g * len + row];
      s0 += pr[(0 * G + g) * len + row];
      s1 += pr[(1 * G + g) * len + row];
      s2 += pr[(2 * G + g) * len + row];
    }
    gD[0 * len + row] += maj;
    gD[3 * len + row] += maj;
    gD[5 * len + row] += maj;
    gR[0 * len + row] += maj * gX[0 * len + row] - s0;
    gR[1 * len + row] += maj * gX[1 * len + row] - s1;
    gR[2 * len + row] += maj * gX[2 * len + row] - s2;
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t >= Lr) return;  // the solve uses the first Lr threads only (no barrier below involves the rest)
  float Da[6], Ua[9], ra[3], Db[6], Ub[9], rb[3];
  const int ka = 2 * t, kb = 2 * t + 1;
  if (ka < len) {
#pragma unroll
    for (int q = 0; q < 6; q++) Da[q] = gD[q * len + ka];
#pragma unroll
    for (int q = 0; q < 9; q++) Ua[q] = gU[q * len + ka];
#pragma unroll
    for (int q = 0; q < 3; q++) ra[q] = gR[q * len + ka];
  } else {
    set_identity(Da, Ua, ra);
  }
  if (kb < len) {
#pragma unroll
    for (int q = 0; q < 6; q++) Db[q] = gD[q * len + kb];
#pragma unroll
    for (int q = 0; q < 9; q++) Ub[q] = gU[q * len + kb];
#pragma unroll
    for (int q = 0; q < 3; q++) rb[q] = gR[q * len + kb];
  } else {
    set_identity(Db, Ub, rb);
  }
  float xa[3], xb[3];
  line_solve(cr, Lr, t, len, Da, Ua, ra, Db, Ub, rb, xa, xb);
  if (ka < len && !gF[ka]) {
    int n, nb, no, i, j;
    row_index<FAM, DUAL>(L, ka, n, nb, no, i, j);
    store_x(S, n, no, xa, DUAL);
  }
  if (kb < len && !gF[kb]) {
    int n, nb, no, i, j;
    row_index<FAM, DUAL>(L, kb, n, nb, no, i, j);
    store_x(S, n, no, xb, DUAL);
  }
}

// Split variant, first kernel: one thread per row of the pass's lines (blockIdx.y the line, x the
// 256-row chunk), so that the whole chip gathers with full occupancy and no line's workgroup waits
// on another's CU; writes the row's D, U and r to the rows planes (line slot * stride + row).
template <bool FINE, int FAM, bool DUAL, class RDD>
__global__ __launch_bounds__(256) void k_line_rows(SmoothArgs S, RDD rdd, float* rows, size_t cap, int stride) {
  if (S.gate && *S.gate) return;
  const auto rd = rdd.live();
  const int slot = blockIdx.y;
  const LineGeo L = line_geo_at<FAM, DUAL>(S, slot);
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= L.ln.z) return;
  int n, nb, no, i, j;
  row_index<FAM, DUAL>(L, k, n, nb, no, i, j);
  float D[6], U[9], r[3];
  bool f;
  gather_row<FINE, FAM, DUAL>(S, rd, n, nb, i, j, L.nx, L.ny, D, U, r, f);
  float* o = rows + (size_t)slot * stride + k;
#pragma unroll
  for (int q = 0; q < 6; q++) o[q * cap] = D[q];
#pragma unroll
  for (int q = 0; q < 9; q++) o[(6 + q) * cap] = U[q];
#pragma unroll
  for (int q = 0; q < 3; q++) o[(15 + q) * cap] = r[q];
}

// Split variant, second kernel: the line solve of k_smooth_direct over the rows k_line_rows wrote.
template <bool FINE, int FAM, bool DUAL>
__global__ __launch_bounds__(512, 4) void k_smooth_rows(SmoothArgs S, const float* rows, size_t cap, int stride) {
  if (S.gate && *S.gate) return;
  extern __shared__ __align__(16) float sm[];
  const int slot = xcd_block(blockIdx.x, gridDim.x);
  const LineGeo L = line_geo_at<FAM, DUAL>(S, slot);
  const int len = L.ln.z;
  const int t = threadIdx.x;
  const int ka = 2 * t, kb = 2 * t + 1;
  int na, nba, noa, ia, ja, nb, nbb, nob, ib, jb;
  row_index<FAM, DUAL>(L, ka, na, nba, noa, ia, ja);
  row_index<FAM, DUAL>(L, kb, nb, nbb, nob, ib, jb);
  float Da[6], Ua[9], ra[3], Db[6], Ub[9], rb[3];
  // rows ka and kb are adjacent: one 8-byte load per plane (stride and cap are even)
  const float* o = rows + (size_t)slot * stride;
  float v[18][2];
#pragma unroll
  for (int q = 0; q < 18; q++) {
    const float2 p = ka < len ? *reinterpret_cast<const float2*>(o + q * cap + ka) : make_float2(0.f, 0.f);
    v[q][0] = p.x;
    v[q][1] = p.y;
  }
  if (ka < len) {
#pragma unroll
    for (int q = 0; q < 6; q++) Da[q] = v[q][0];
#pragma unroll
    for (int q = 0; q < 9; q++) Ua[q] = v[6 + q][0];
#pragma unroll
    for (int q = 0; q < 3; q++) ra[q] = v[15 + q][0];
  } else {
    set_identity(Da, Ua, ra);
  }
  if (kb < len) {
#pragma unroll
    for (int q = 0; q < 6; q++) Db[q] = v[q][1];
#pragma unroll
    for (int q = 0; q < 9; q++) Ub[q] = v[6 + q][1];
#pragma unroll
    for (int q = 0; q < 3; q++) rb[q] = v[15 + q][1];
  } else {
    set_identity(Db, Ub, rb);
  }
  const bool fa = ka >= len || S.fixed[na] != 0, fb = kb >= len || S.fixed[nb] != 0;
  float xa[3], xb[3];
  line_solve(sm, blockDim.x, t, len, Da, Ua, ra, Db, Ub, rb, xa, xb);
  if (!fa) store_x(S, na, noa, xa, DUAL);
  if (!fb) store_x(S, nb, nob, xb, DUAL);
}

template <bool FINE, int FAM, int G>
size_t staged_shm(int len) {
  int R2 = (len + 1) / 2;
  int Lr = ((R2 + 63) / 64) * 64;
  return sizeof(float) * ((size_t)kSolveFloats * Lr + (size_t)(6 + 9 + 3 + 3 + 4 * G) * len) + (size_t)len + 16;
}

// LDS a workgroup may use (the staged kernels' attribute is raised to it for the autotuner)
constexpr int kMaxLds = 160 * 1024;

// variant: -1 the default rule below, 0 the direct kernel, 4 or 8 the staged kernel with that gather
// width (bd_autotune picks per level and family)
template <bool FINE, int FAM, bool DUAL, class RDD>
void launch_one(const SmoothArgs& S, RDD rdd, int count, int len, hipStream_t st, int variant, float* rows, size_t cap) {
  if (variant == kSplitVariant) {
    const int stride = (len + 1) & ~1;  // even, so that a row pair is one aligned float2
    if (!rows || (size_t)count * stride > cap) throw std::runtime_error("smoother: split variant without room for its rows");
    k_line_rows<FINE, FAM, DUAL, RDD><<<dim3((len + 255) / 256, count), 256, 0, st>>>(S, rdd, rows, cap, stride);
    const int block = ((len + 1) / 2 + 63) / 64 * 64;
    k_smooth_rows<FINE, FAM, DUAL><<<count, block, (size_t)kSolveFloats * block * sizeof(float), st>>>(S, rows, cap, stride);
    return;
  }
  if (variant == 4 || variant == 8) {
    const int G = variant;
    const int R2 = (len + 1) / 2;
    const int threads = std::max(std::min(1024, ((len * G + 63) / 64) * 64), ((R2 + 63) / 64) * 64);
    if (G == 8)
      k_smooth_staged<FINE, FAM, 8, DUAL, RDD><<<count, threads, staged_shm<FINE, FAM, 8>(len), st>>>(S, rdd);
    else
      k_smooth_staged<FINE, FAM, 4, DUAL, RDD><<<count, threads, staged_shm<FINE, FAM, 4>(len), st>>>(S, rdd);
    return;
  }
  if (variant < 0 && len <= kStagedMaxLen) {
    // widest gather that fits 1024 threads: couplings per thread ~ (FINE ? 12 : 24) / G
    const int G = len <= 128 ? 8 : 4;
    int threads = std::min(1024, ((len * G + 63) / 64) * 64);
    const int R2 = (len + 1) / 2;
    threads = std::max(threads, ((R2 + 63) / 64) * 64);
    if (G == 8)
      k_smooth_staged<FINE, FAM, 8, DUAL, RDD><<<count, threads, staged_shm<FINE, FAM, 8>(len), st>>>(S, rdd);
    else
      k_smooth_staged<FINE, FAM, 4, DUAL, RDD><<<count, threads, staged_shm<FINE, FAM, 4>(len), st>>>(S, rdd);
    return;
  }
  int R2 = (len + 1) / 2;
  int block = ((R2 + 63) / 64) * 64;
  size_t shm = (size_t)kSolveFloats * block * sizeof(float);
  k_smooth_direct<FINE, FAM, DUAL, RDD><<<count, block, shm, st>>>(S, rdd);
}

template <bool F, int A, bool D, class R>
void set_direct_attr(int shm) {
  BD_HIP(hipFuncSetAttribute((const void*)k_smooth_direct<F, A, D, R>, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
}
template <bool F, int A, int G, bool D, class R>
void set_staged_attr(int shm) {
  BD_HIP(hipFuncSetAttribute((const void*)k_smooth_staged<F, A, G, D, R>, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
}
template <bool F, int A, bool D>
void set_rows_attr(int shm) {
  BD_HIP(hipFuncSetAttribute((const void*)k_smooth_rows<F, A, D>, hipFuncAttributeMaxDynamicSharedMemorySize, shm));
}
template <bool D, class R>
void set_attrs(int shm, int shm_staged) {
  set_direct_attr<true, 0, D, R>(shm);
  set_direct_attr<true, 1, D, R>(shm);
  set_direct_attr<false, 0, D, R>(shm);
  set_direct_attr<false, 1, D, R>(shm);
  set_staged_attr<true, 0, 4, D, R>(shm_staged);
  set_staged_attr<true, 1, 4, D, R>(shm_staged);
  set_staged_attr<true, 0, 8, D, R>(shm_staged);
  set_staged_attr<true, 1, 8, D, R>(shm_staged);
  set_staged_attr<false, 0, 4, D, R>(shm_staged);
  set_staged_attr<false, 1, 4, D, R>(shm_staged);
  set_staged_attr<false, 0, 8, D, R>(shm_staged);
  set_staged_attr<false, 1, 8, D, R>(shm_staged);
}

}  // namespace

void smooth_set_attributes() {
  int shm = kSolveFloats * ((BD_MAX_LINE / 2 + 63) / 64 * 64) * (int)sizeof(float);
  set_attrs<false, OpF32>(shm, kMaxLds);
  set_attrs<true, OpF32>(shm, kMaxLds);
  set_attrs<false, OpF16Dev>(shm, kMaxLds);
  set_attrs<true, OpF16Dev>(shm, kMaxLds);
  set_rows_attr<true, 0, false>(shm);
  set_rows_attr<true, 1, false>(shm);
  set_rows_attr<false, 0, false>(shm);
  set_rows_attr<false, 1, false>(shm);
  set_rows_attr<true, 0, true>(shm);
  set_rows_attr<true, 1, true>(shm);
  set_rows_attr<false, 0, true>(shm);
  set_rows_attr<false, 1, true>(shm);
}

namespace {

// one colour pass of level k with the given variant
void launch_colour(Solver& s, int k, int c, int variant) {
  Level& L = s.lv[k];
  int count = L.color_start[c + 1] - L.color_start[c];
  if (count == 0) return;
  const int fam = c < L.ncol ? 1 : 0;
  int maxlen = 0;
  for (int l = 0; l < L.geo.nl; l++) maxlen = std::max(maxlen, fam == 1 ? L.geo.ny[l] : L.geo.nx[l]);
  const bool colmajor = L.dual && fam == 1;
  SmoothArgs S{colmajor ? L.opT : L.op, colmajor ? L.fixedT : L.fixed, L.Np, L.lines + L.color_start[c],
               colmajor ? L.xT : L.x, colmajor ? L.bT : L.b, L.dual ? (colmajor ? L.x : L.xT) : nullptr, &s.scal->done,
               L.geo};
  // each node of the colour's lines: its full row (fine 58 floats; coarse 150 values, 4 B each or 2 B
  // as FP16 pairs) + b, x read, x write (dual-layout levels write x twice)
  double nodes = (double)count * maxlen;
  const double row = L.fine ? 268.0 : (L.oph ? 336.0 : 636.0);
  ProfScope ps(s, L.fine ? PC_SMOOTH_FINE : PC_SMOOTH_COARSE, nodes * (row + (L.dual ? 12.0 : 0.0)));
  auto go = [&](auto rdd) {
    using R = decltype(rdd);
    if (L.dual) {
      if (L.fine) {
        if (fam == 0) launch_one<true, 0, true, R>(S, rdd, count, maxlen, s.stream, variant, L.rows, L.rows_cap);
        else launch_one<true, 1, true, R>(S, rdd, count, maxlen, s.stream, variant, L.rows, L.rows_cap);
      } else {
        if (fam == 0) launch_one<false, 0, true, R>(S, rdd, count, maxlen, s.stream, variant, L.rows, L.rows_cap);
        else launch_one<false, 1, true, R>(S, rdd, count, maxlen, s.stream, variant, L.rows, L.rows_cap);
      }
    } else {
      if (L.fine) {
        if (fam == 0) launch_one<true, 0, false, R>(S, rdd, count, maxlen, s.stream, variant, L.rows, L.rows_cap);
        else launch_one<true, 1, false, R>(S, rdd, count, maxlen, s.stream, variant, L.rows, L.rows_cap);
      } else {
        if (fam == 0) launch_one<false, 0, false, R>(S, rdd, count, maxlen, s.stream, variant, L.rows, L.rows_cap);
        else launch_one<false, 1, false, R>(S, rdd, count, maxlen, s.stream, variant, L.rows, L.rows_cap);
      }
    }
  };
  // the operator in the pass's layout, FP32 or FP16 pairs
  if (L.oph)
    go(OpF16Dev{colmajor ? L.ophT : L.oph, L.Np, L.hs});
  else
    go(OpF32{colmajor ? L.opT : L.op, L.Np});
}

// the variants a line of `len` rows can run: the direct kernel, the split one where the level keeps
// a rows buffer, and the staged ones whose LDS fits
std::vector<int> variants_for(bool fine, int len, bool has_rows) {
  std::vector<int> v{0};
  if (has_rows) v.push_back(kSplitVariant);
  if ((fine ? staged_shm<true, 0, 4>(len) : staged_shm<false, 0, 4>(len)) <= (size_t)kMaxLds) v.push_back(4);
  if ((fine ? staged_shm<true, 0, 8>(len) : staged_shm<false, 0, 8>(len)) <= (size_t)kMaxLds) v.push_back(8);
  return v;
}

}  // namespace

// The kernel variants a level's passes of one family can run.
std::vector<int> smooth_variants(Solver& s, int k, int fam) {
  Level& L = s.lv[k];
  int maxlen = 0;
  for (int l = 0; l < L.geo.nl; l++) maxlen = std::max(maxlen, fam == 1 ? L.geo.ny[l] : L.geo.nx[l]);
  return variants_for(L.fine, maxlen, L.rows != nullptr);
}

// The rows buffer of the split variant: 18 planes of (most lines of one colour) x (longest line).
void smooth_alloc_rows(Solver& s, Level& L) {
  (void)s;
  int most = 0;
  for (int c = 0; c < 2 * L.ncol; c++) most = std::max(most, L.color_start[c + 1] - L.color_start[c]);
  L.rows_cap = (size_t)most * ((L.max_len + 1) & ~1);
  BD_HIP(hipMalloc(&L.rows, 18 * L.rows_cap * sizeof(float)));
}

// Smooth all colours of level k (v-line colours first, then u-line colours; reversed for the
// post-smoothing sweep so the V-cycle stays symmetric).
void launch_smooth(Solver& s, int k, int reverse) {
  Level& L = s.lv[k];
  const int nc = 2 * L.ncol;
  for (int q = 0; q < nc; q++) {
    int c = reverse ? nc - 1 - q : q;
    launch_colour(s, k, c, L.variant[c < L.ncol ? 1 : 0]);
  }
}

}  // namespace bd

// Device code of the multigrid transfers shared by the V-cycle kernels (multigrid.hip) and the
// one-workgroup solver (wgsolve.hip): operator rows, the branch-free V-cycle residual row, the
// coarse-to-fine index map and bracket, restriction and prolongation at one node (lattice.ts:13-41,
// newtonMultigrid.ts:287-387, 495-528).
#pragma once
#include "bd_internal.h"
#include "oprd.h"

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

// y = A x for row n in the V-cycle, through an operator reader (FP32 or FP16 planes). Branch-free:
// absent neighbours read row n itself and are masked, so every load of the row issues at once. The
// couplings are summed in the order of for_couplings.
template <bool FINE, class R>
__device__ __forceinline__ void vc_apply_row(const R& rd, const Geo& g, const float* x, int Np, int n, float y[3]) {
  int l, i, j;
  node_ij(g, n, l, i, j);
  const int nx = g.nx[l], ny = g.ny[l];
  float c[6], xv[3] = {x[n], x[Np + n], x[2 * Np + n]};
  rd.blk(F_C, n, c);
  sym_mul(c, xv, y);
  auto one = [&](bool in, int base, int src, int m) {
    float b[6], t[3];
    rd.blk(base, in ? src : n, b);
    const int mm = in ? m : n;
    const float xm[3] = {x[mm], x[Np + mm], x[2 * Np + mm]};
    sym_mul(b, xm, t);
    const float w = in ? 1.f : 0.f;
    y[0] += w * t[0];
    y[1] += w * t[1];
    y[2] += w * t[2];
  };
  auto ones = [&](bool in, int plane, int src, int m) {
    const float s = in ? rd.scalar(plane, src) : 0.f;
    const int mm = in ? m : n;
    y[0] += s * x[mm];
    y[1] += s * x[Np + mm];
    y[2] += s * x[2 * Np + mm];
  };
  if (FINE) {
    const bool e = i + 1 < nx, w = i > 0, no = j + 1 < ny, so = j > 0;
    one(e, F_E, n, n + 1);
    one(w, F_E, n - 1, n - 1);
    one(no, F_N, n, n + nx);
    one(so, F_N, n - nx, n - nx);
    one(e && no, F_NE, n, n + nx + 1);
    one(w && so, F_NE, n - nx - 1, n - nx - 1);
    one(w && no, F_NW, n, n + nx - 1);
    one(e && so, F_NW, n - nx + 1, n - nx + 1);
    ones(i + 2 < nx, F_EE, n, n + 2);
    ones(i >= 2, F_EE, n - 2, n - 2);
    ones(j + 2 < ny, F_NN, n, n + 2 * nx);
    ones(j >= 2, F_NN, n - 2 * nx, n - 2 * nx);
  } else {
    // fully unrolled so that the 24 couplings' loads issue together
#pragma clang loop unroll(full)
    for (int k = 1; k < kCoarseBlocks; k++) {
      const int di = coarse_di(k), dj = coarse_dj(k);
      const int mf = n + dj * nx + di, mb = n - dj * nx - di;
      one(i + di >= 0 && i + di < nx && j + dj < ny, 6 * k, n, mf);
      one(i - di >= 0 && i - di < nx && j - dj >= 0, 6 * k, mb, mb);
    }
  }
}

// at most 4 waves per SIMD: the register room lets a row's couplings load together (at the default
// occupancy target the scheduler waits for each coupling before loading the next)

__device__ __forceinline__ int ccount(int n) { return n / 2 + 1; }  // ceil((n-1)/2) + 1
__device__ __forceinline__ int c2f(int I, int n) { return 2 * I < n - 1 ? 2 * I : n - 1; }

// bracket's first coarse index and count without the weights
__device__ __forceinline__ int bracket_lo(int i, int n, int& lo) {
  const int nc = ccount(n);
  lo = (i >> 1) < nc - 1 ? (i >> 1) : nc - 1;
  if (c2f(lo, n) > i) lo--;
  return c2f(lo, n) == i ? 1 : 2;
}

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
  float* xcT;      // column-major coarse solution (dual-layout coarse level) or null
  float* xfT;      // column-major fine solution (dual-layout fine level) or null
  float* bcT = nullptr;  // column-major coarse rhs (dual-layout coarse level) or null
};

// coarse correction at fine node f = (l, i, j): the up to 2 x 2 bracketing coarse nodes,
// branch-free (absent terms read the first one with weight 0)
__device__ __forceinline__ void prolong_sum(const XferArgs& X, int f, int l, int i, int j, float s[3]) {
  const int nx = X.a.geo.nx[l], ny = X.a.geo.ny[l];
  const int cnx = X.cgeo.nx[l], coff = X.cgeo.off[l];
  const int Np = X.a.Np, Npc = X.Npc;
  int loj, loi;
  const int nj = bracket_lo(j, ny, loj), ni = bracket_lo(i, nx, loi);
  s[0] = s[1] = s[2] = 0.f;
#pragma unroll
  for (int ej = 0; ej < 2; ej++)
#pragma unroll
    for (int ei = 0; ei < 2; ei++) {
      const bool use = ej < nj && ei < ni;
      const float wv = X.pw[(use ? ej * ni + ei : 0) * Np + f];
      const float w = use ? wv : 0.f;
      const int c = coff + (loj + (use ? ej : 0)) * cnx + loi + (use ? ei : 0);
      s[0] += w * X.xc[c];
      s[1] += w * X.xc[Npc + c];
      s[2] += w * X.xc[2 * Npc + c];
    }
}

// restriction at coarse node I: the masked weights of the 3 x 3 fine nodes around its fine position
// times the fine residual
__device__ __forceinline__ void restrict_node(const XferArgs& X, int I, float s[3]) {
  int l, Ii, Ij;
  node_ij(X.cgeo, I, l, Ii, Ij);
  const int nx = X.a.geo.nx[l], ny = X.a.geo.ny[l], foff = X.a.geo.off[l];
  const int Np = X.a.Np;
  s[0] = s[1] = s[2] = 0.f;
  {
    // the 3 x 3 fine nodes around the coarse node's fine position, branch-free: an absent node, a
    // fixed one or a weight plane that does not reach this coarse node reads row a0 and adds 0
    const int cfi = c2f(Ii, nx), cfj = c2f(Ij, ny);
    const bool cfree = !X.cfixed[I];
    const int a0 = foff + cfj * nx + cfi;
#pragma unroll
    for (int dj = -1; dj <= 1; dj++)
#pragma unroll
      for (int di = -1; di <= 1; di++) {
        const int ai = cfi + di, aj = cfj + dj;
        const bool in = ai >= 0 && ai < nx && aj >= 0 && aj < ny;
        const int a = in ? foff + aj * nx + ai : a0;
        int loj, loi;
        const int nj = bracket_lo(in ? aj : cfj, ny, loj), ni = bracket_lo(in ? ai : cfi, nx, loi);
        const int ej = Ij - loj, ei = Ii - loi;
        const bool reach = ej >= 0 && ej < nj && ei >= 0 && ei < ni;
        const float wv = X.pw[(reach ? ej * ni + ei : 0) * Np + a];
        const uint8_t fx = X.a.fixed[a];
        const float w = (cfree & in & reach & (fx == 0)) ? wv : 0.f;
        s[0] += w * X.r[a];
        s[1] += w * X.r[Np + a];
        s[2] += w * X.r[2 * Np + a];
      }
  }
}

// y += A d at fine row n with d in FP64 and FP64 accumulation over the FP32 operator (the true
// residual of residual replacement); couplings in the order of the FP32 matvec
__device__ __forceinline__ void fine_apply_f64(const float* op, int Np, const Geo& geo, const double* delta, int n,
                                               double y[3]) {
  int l = geo_layer(geo, n);
  int r0 = n - geo.off[l], nx = geo.nx[l], ny = geo.ny[l];
  int j = r0 / nx, i = r0 - j * nx;
  const double* d = delta;
  auto acc = [&](int base, int row, int m) {
    double s0 = op[(base + 0) * Np + row], s1 = op[(base + 1) * Np + row], s2 = op[(base + 2) * Np + row];
    double s3 = op[(base + 3) * Np + row], s4 = op[(base + 4) * Np + row], s5 = op[(base + 5) * Np + row];
    double x0 = d[m], x1 = d[Np + m], x2 = d[2 * Np + m];
    y[0] += s0 * x0 + s1 * x1 + s2 * x2;
    y[1] += s1 * x0 + s3 * x1 + s4 * x2;
    y[2] += s2 * x0 + s4 * x1 + s5 * x2;
  };
  auto accs = [&](double s, int m) {
    y[0] += s * d[m];
    y[1] += s * d[Np + m];
    y[2] += s * d[2 * Np + m];
  };
  acc(F_C, n, n);
  const bool e = i + 1 < nx, w = i > 0, no = j + 1 < ny, so = j > 0;
  if (e) acc(F_E, n, n + 1);
  if (w) acc(F_E, n - 1, n - 1);
  if (no) acc(F_N, n, n + nx);
  if (so) acc(F_N, n - nx, n - nx);
  if (e && no) acc(F_NE, n, n + nx + 1);
  if (w && so) acc(F_NE, n - nx - 1, n - nx - 1);
  if (w && no) acc(F_NW, n, n + nx - 1);
  if (e && so) acc(F_NW, n - nx + 1, n - nx + 1);
  if (i + 2 < nx) accs(op[F_EE * Np + n], n + 2);
  if (i >= 2) accs(op[F_EE * Np + n - 2], n - 2);
  if (j + 2 < ny) accs(op[F_NN * Np + n], n + 2 * nx);
  if (j >= 2) accs(op[F_NN * Np + n - 2 * nx], n - 2 * nx);
}

}  // namespace

}  // namespace bd

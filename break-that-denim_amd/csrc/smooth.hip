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
#include "smooth_dev.h"

namespace bd {

namespace {

// Long lines: each thread gathers its two rows straight into registers.
template <bool FINE, int FAM, bool DUAL, class RDD>
__global__ __launch_bounds__(512, 4) void k_smooth_direct(SmoothArgs S, RDD rdd) {
  const PassHead h = pass_head(S, xcd_block(blockIdx.x, gridDim.x));
  if (h.gated || h.ln.z <= 0) return;
  const auto rd = rdd.live();
  extern __shared__ __align__(16) float sm[];
  const LineGeo L = line_geo_of<FAM, DUAL>(S, h.ln);
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
  const PassHead h = pass_head(S, xcd_block(blockIdx.x, gridDim.x));
  if (h.gated || h.ln.z <= 0) return;
  const auto rd = rdd.live();
  extern __shared__ __align__(16) float sm[];
  const LineGeo L = line_geo_of<FAM, DUAL>(S, h.ln);
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
  // (line, 256-row chunk) of this block, consecutive lines on one XCD so that the couplings and x
  // a line shares with its neighbours are read through one L2
  const int b = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int slot = b / gridDim.x;
  const PassHead h = pass_head(S, slot);
  if (h.gated || h.ln.z <= 0) return;
  const auto rd = rdd.live();
  const LineGeo L = line_geo_of<FAM, DUAL>(S, h.ln);
  const int k = (b - slot * gridDim.x) * 256 + threadIdx.x;
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
  extern __shared__ __align__(16) float sm[];
  const int slot = xcd_block(blockIdx.x, gridDim.x);
  const PassHead h = pass_head(S, slot);
  if (h.gated || h.ln.z <= 0) return;
  const LineGeo L = line_geo_of<FAM, DUAL>(S, h.ln);
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

// Split variant with four rows per thread, first kernel: as k_line_rows, but only each row's
// right-hand side r (3 planes): the line blocks D and U do not change with the iterate, so the solve
// kernel reads them from the operator itself and the round trip of 15 of the 18 floats is saved.
template <bool FINE, int FAM, bool DUAL, class RDD>
__global__ __launch_bounds__(256) void k_line_r(SmoothArgs S, RDD rdd, float* rows, size_t cap, int stride) {
  // (line, 256-row chunk) of this block, consecutive lines on one XCD so that the couplings and x
  // a line shares with its neighbours are read through one L2
  const int b = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int slot = b / gridDim.x;
  const PassHead h = pass_head(S, slot);
  if (h.gated || h.ln.z <= 0) return;
  const auto rd = rdd.live();
  const LineGeo L = line_geo_of<FAM, DUAL>(S, h.ln);
  const int k = (b - slot * gridDim.x) * 256 + threadIdx.x;
  if (k >= L.ln.z) return;
  int n, nb, no, i, j;
  row_index<FAM, DUAL>(L, k, n, nb, no, i, j);
  float D[6], U[9], r[3];
  bool f;
  gather_row<FINE, FAM, DUAL, decltype(rd), kGatherR>(S, rd, n, nb, i, j, L.nx, L.ny, D, U, r, f);
  float* o = rows + (size_t)slot * stride + k;
#pragma unroll
  for (int q = 0; q < 3; q++) o[q * cap] = r[q];
}

// The line blocks D (6) and U (9) of every row of the lines of one family (blockIdx.y the line within
// the family, x the 256-row chunk), as gather_row assembles them, to planes at (line0 + line) *
// stride + row of du. Launched when the level's operator is final, outside PCG (no gate).
template <bool FINE, int FAM, bool DUAL, class RDD>
__global__ __launch_bounds__(256) void k_line_du(SmoothArgs S, RDD rdd, float* du, size_t cap, int stride, int line0) {
  const auto rd = rdd.live();
  const int slot = blockIdx.y;
  const LineGeo L = line_geo_at<FAM, DUAL>(S, slot);
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= L.ln.z) return;
  int n, nb, no, i, j;
  row_index<FAM, DUAL>(L, k, n, nb, no, i, j);
  float D[6], U[9], r[3];
  bool f;
  gather_row<FINE, FAM, DUAL, decltype(rd), kGatherDU>(S, rd, n, nb, i, j, L.nx, L.ny, D, U, r, f);
  float* o = du + (size_t)(line0 + slot) * stride + k;
#pragma unroll
  for (int q = 0; q < 6; q++) o[q * cap] = D[q];
#pragma unroll
  for (int q = 0; q < 9; q++) o[(6 + q) * cap] = U[q];
}

// Split variant with four rows per thread, second kernel: r from k_line_r and D, U from the level's
// line blocks (k_line_du), for each (stride a multiple of 4, so that a thread's four rows of a plane
// are one aligned float4), solved by line_solve4.
template <bool FINE, int FAM, bool DUAL>
__global__ __launch_bounds__(256) void k_smooth_rows4(SmoothArgs S, const float* rows, size_t cap, int stride,
                                                      const float* du, size_t du_cap, int du_stride, int line0) {
  extern __shared__ __align__(16) float sm[];
  const int slot = xcd_block(blockIdx.x, gridDim.x);
  const PassHead h = pass_head(S, slot);
  const int t = threadIdx.x;
  const int k0 = 4 * t;
  const float* o = rows + (size_t)slot * stride;
  const float* od = du + (size_t)(line0 + slot) * du_stride;
  // the planes' loads do not depend on the line entry: issued with the gate's, at indices clamped
  // into the line's stride, and masked once the length is known
  const int kd = min(k0, du_stride - 4), kr = min(k0, stride - 4);
  float v[18][4];
#pragma unroll
  for (int q = 0; q < 18; q++) {
    const float* src = q < 15 ? od + q * du_cap + kd : o + (q - 15) * cap + kr;
    const float4 p = *reinterpret_cast<const float4*>(src);
    v[q][0] = p.x;
    v[q][1] = p.y;
    v[q][2] = p.z;
    v[q][3] = p.w;
  }
  if (h.gated || h.ln.z <= 0) return;
  const LineGeo L = line_geo_of<FAM, DUAL>(S, h.ln);
  const int len = L.ln.z;
  int nn[4], no[4];
  uint8_t fx[4];
  line_rows_fixed<FAM, DUAL>(S, L, k0, len, nn, no, fx);
  float D[4][6], U[4][9], r[4][3], x[4][3];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    if (k0 + m < len) {
#pragma unroll
      for (int q = 0; q < 6; q++) D[m][q] = v[q][m];
#pragma unroll
      for (int q = 0; q < 9; q++) U[m][q] = v[6 + q][m];
#pragma unroll
      for (int q = 0; q < 3; q++) r[m][q] = v[15 + q][m];
    } else {
      set_identity(D[m], U[m], r[m]);
    }
  }
  line_solve4(sm, blockDim.x, t, len, D, U, r, x);
#pragma unroll
  for (int m = 0; m < 4; m++)
    if (k0 + m < len && !fx[m]) store_x(S, nn[m], no[m], x[m], DUAL);
}

// The factorization of every line of one family (blockIdx.x the line within the family, four rows per
// thread) from its line blocks (k_line_du), to the factor planes at the same positions. Launched with
// the line blocks, when the level's operator is final (no gate).
template <int FAM, bool DUAL>
__global__ __launch_bounds__(256) void k_line_factor(SmoothArgs S, const float* du, float* fac, size_t cap, int stride,
                                                     int line0) {
  extern __shared__ __align__(16) float sm[];
  const LineGeo L = line_geo_at<FAM, DUAL>(S, blockIdx.x);
  const int len = L.ln.z;
  const int t = threadIdx.x;
  const int k0 = 4 * t;
  const float* od = du + (size_t)(line0 + blockIdx.x) * stride;
  float* of = fac + (size_t)(line0 + blockIdx.x) * stride;
  float v[15][4];
#pragma unroll
  for (int q = 0; q < 15; q++) {
    const float4 p = k0 < len ? *reinterpret_cast<const float4*>(od + q * cap + k0) : make_float4(0.f, 0.f, 0.f, 0.f);
    v[q][0] = p.x;
    v[q][1] = p.y;
    v[q][2] = p.z;
    v[q][3] = p.w;
  }
  float D[4][6], U[4][9], r[3];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    if (k0 + m < len) {
#pragma unroll
      for (int q = 0; q < 6; q++) D[m][q] = v[q][m];
#pragma unroll
      for (int q = 0; q < 9; q++) U[m][q] = v[6 + q][m];
    } else {
      set_identity(D[m], U[m], r);
    }
  }
  line_factor4(sm, of, cap, t, len, D, U);
}

// Factored split variant, second kernel: r from k_line_r and the rows' factors (k_line_factor), four
// rows per thread, solved by line_solve_factored: the pass moves r and x through the elimination
// kept from the factorization instead of eliminating the blocks again.
template <bool FINE, int FAM, bool DUAL>
__global__ __launch_bounds__(256) void k_smooth_factored(SmoothArgs S, const float* rows, size_t cap, int stride,
                                                         const float* fac, size_t fcap, int fstride, int line0) {
  extern __shared__ __align__(16) float4 sm4[];
  const int slot = xcd_block(blockIdx.x, gridDim.x);
  const PassHead h = pass_head(S, slot);
  const int t = threadIdx.x;
  const int k0 = 4 * t;
  const float* o = rows + (size_t)slot * stride;
  const float* of = fac + (size_t)(line0 + slot) * fstride;
  // the planes' loads do not depend on the line entry: issued with the gate's, at indices clamped
  // into the line's stride, and masked once the length is known
  const int kf = min(k0, fstride - 4), kr = min(k0, stride - 4);
  float F[kFacPlanes][4], r[4][3], x[4][3];
#pragma unroll
  for (int q = 0; q < kFacPlanes; q++) {
    const float4 p = *reinterpret_cast<const float4*>(of + q * fcap + kf);
    F[q][0] = p.x;
    F[q][1] = p.y;
    F[q][2] = p.z;
    F[q][3] = p.w;
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const float4 p = *reinterpret_cast<const float4*>(o + c * cap + kr);
    r[0][c] = p.x;
    r[1][c] = p.y;
    r[2][c] = p.z;
    r[3][c] = p.w;
  }
  if (h.gated || h.ln.z <= 0) return;
  const LineGeo L = line_geo_of<FAM, DUAL>(S, h.ln);
  const int len = L.ln.z;
  int nn[4], no[4];
  uint8_t fx[4];
  line_rows_fixed<FAM, DUAL>(S, L, k0, len, nn, no, fx);
  // rows past the line: no factor and no right-hand side, so they contribute nothing and solve to 0
#pragma unroll
  for (int m = 0; m < 4; m++) {
    if (k0 + m >= len) {
#pragma unroll
      for (int q = 0; q < kFacPlanes; q++) F[q][m] = 0.f;
      r[m][0] = r[m][1] = r[m][2] = 0.f;
    }
  }
  line_solve_factored(sm4, blockDim.x, t, len, F, r, x);
#pragma unroll
  for (int m = 0; m < 4; m++)
    if (k0 + m < len && !fx[m]) store_x(S, nn[m], no[m], x[m], DUAL);
}

template <bool FINE, int FAM, int G>
size_t staged_shm(int len) {
  int R2 = (len + 1) / 2;
  int Lr = ((R2 + 63) / 64) * 64;
  return sizeof(float) * ((size_t)kSolveFloats * Lr + (size_t)(6 + 9 + 3 + 3 + 4 * G) * len) + (size_t)len + 16;
}

// Threads of a staged pass: one wave per gather item (64 consecutive rows x one of G coupling parts),
// so that the whole gather is one round of loads, up to 1024 threads.
inline int staged_threads(int len, int G) { return std::min(1024, ((len + 63) / 64) * G * 64); }

// LDS a workgroup may use (the staged kernels' attribute is raised to it for the autotuner)
constexpr int kMaxLds = 160 * 1024;

// variant: -1 the default rule below, 0 the direct kernel, 4 or 8 the staged kernel with that gather
// width (bd_autotune picks per level and family)
template <bool FINE, int FAM, bool DUAL, class RDD>
void launch_one(const SmoothArgs& S, RDD rdd, int count, int len, hipStream_t st, int variant, const Level& Lv, int line0) {
  float* rows = Lv.rows;
  const size_t cap = Lv.rows_cap;
  if (variant == kFactorVariant) {
    const int stride = (len + 3) & ~3;
    if (!rows || (size_t)count * stride > cap || cap % 4 != 0 || !Lv.fac || Lv.du_stride < stride)
      throw std::runtime_error("smoother: factored variant without its factors");
    k_line_r<FINE, FAM, DUAL, RDD><<<dim3((len + 255) / 256, count), 256, 0, st>>>(S, rdd, rows, cap, stride);
    const int block = ((len + 3) / 4 + 63) / 64 * 64;
    k_smooth_factored<FINE, FAM, DUAL><<<count, block, (size_t)6 * block * sizeof(float4), st>>>(
        S, rows, cap, stride, Lv.fac, Lv.du_cap, Lv.du_stride, line0);
    return;
  }
  if (variant == kSplitVariant || variant == kSplit4Variant) {
    const bool four = variant == kSplit4Variant;
    // a thread's rows of a plane are one aligned float2 (two rows) or float4 (four)
    const int stride = four ? (len + 3) & ~3 : (len + 1) & ~1;
    if (!rows || (size_t)count * stride > cap || cap % 4 != 0 || (four && (!Lv.du || Lv.du_stride < stride)))
      throw std::runtime_error("smoother: split variant without room for its rows");
    if (four)
      k_line_r<FINE, FAM, DUAL, RDD><<<dim3((len + 255) / 256, count), 256, 0, st>>>(S, rdd, rows, cap, stride);
    else
      k_line_rows<FINE, FAM, DUAL, RDD><<<dim3((len + 255) / 256, count), 256, 0, st>>>(S, rdd, rows, cap, stride);
    if (four) {
      const int block = ((len + 3) / 4 + 63) / 64 * 64;
      k_smooth_rows4<FINE, FAM, DUAL><<<count, block, (size_t)kSolveFloats * block * sizeof(float), st>>>(
          S, rows, cap, stride, Lv.du, Lv.du_cap, Lv.du_stride, line0);
    } else {
      const int block = ((len + 1) / 2 + 63) / 64 * 64;
      k_smooth_rows<FINE, FAM, DUAL><<<count, block, (size_t)kSolveFloats * block * sizeof(float), st>>>(S, rows, cap, stride);
    }
    return;
  }
  if (variant == 4 || variant == 8) {
    const int G = variant;
    const int R2 = (len + 1) / 2;
    const int threads = std::max(staged_threads(len, G), ((R2 + 63) / 64) * 64);
    if (G == 8)
      k_smooth_staged<FINE, FAM, 8, DUAL, RDD><<<count, threads, staged_shm<FINE, FAM, 8>(len), st>>>(S, rdd);
    else
      k_smooth_staged<FINE, FAM, 4, DUAL, RDD><<<count, threads, staged_shm<FINE, FAM, 4>(len), st>>>(S, rdd);
    return;
  }
  if (variant < 0 && len <= kStagedMaxLen) {
    // widest gather that fits 1024 threads: couplings per thread ~ (FINE ? 12 : 24) / G
    const int G = len <= 128 ? 8 : 4;
    const int R2 = (len + 1) / 2;
    const int threads = std::max(staged_threads(len, G), ((R2 + 63) / 64) * 64);
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

// one colour pass of level k with the given variant; `second` is the second pass of a one-colour
// level's sweep (it reads the pair's other buffer and writes back to x). On a batch_copy level the
// pass writes its own layout only (launch_smooth refreshes the other one).
void launch_colour(Solver& s, int k, int c, int variant, bool second) {
  Level& L = s.lv[k];
  int count = L.color_start[c + 1] - L.color_start[c];
  if (count == 0) return;
  const int fam = c < L.ncol ? 1 : 0;
  int maxlen = 0;
  for (int l = 0; l < L.geo.nl; l++) maxlen = std::max(maxlen, fam == 1 ? L.geo.ny[l] : L.geo.nx[l]);
  const bool colmajor = L.dual && fam == 1;
  // x in the pass's layout and in the other one; a one-colour level's first pass of a sweep writes the
  // pair's second buffers (x2, xT2) from x, the second pass writes x back from them, so that every line
  // reads the lagged iterate of the lines beside it
  float* own = colmajor ? L.xT : L.x;
  float* other = L.dual ? (colmajor ? L.x : L.xT) : nullptr;
  const float* rd_x = own;
  if (L.ncol == 1) {
    float* own2 = colmajor ? L.xT2 : L.x2;
    float* other2 = L.dual ? (colmajor ? L.x2 : L.xT2) : nullptr;
    if (second) {
      rd_x = own2;
    } else {
      own = own2;
      other = other2;
    }
  }
  if (L.batch_copy) other = nullptr;
  // two colours run only the split kernels (the default rule would pick a fused one)
  if (L.ncol == 2 && variant != kSplitVariant && variant != kSplit4Variant) variant = kFactorVariant;
  SmoothArgs S{colmajor ? L.opT : L.op, colmajor ? L.fixedT : L.fixed, L.Np, L.lines + L.color_start[c],
               own, colmajor ? L.bT : L.b, other, &s.scal->done, L.geo, rd_x, L.ncol == 2 ? 1 : 0};
  // each node of the colour's lines: its full row (fine 58 floats; coarse 150 values, 4 B each or 2 B
  // as FP16 pairs) + b, x read, x write (dual-layout levels that copy x per pass write it twice). A
  // one-colour pass holds every line of the family, so the mirrored couplings a row reads are other
  // rows of the same pass: its unique bytes are the operator once (fine 32 planes, coarse 78).
  double nodes = (double)count * maxlen;
  const double row = L.ncol == 1 ? (L.fine ? 164.0 : (L.oph ? 192.0 : 348.0)) : (L.fine ? 268.0 : (L.oph ? 336.0 : 636.0));
  ProfScope ps(s, L.fine ? PC_SMOOTH_FINE : PC_SMOOTH_COARSE, nodes * (row + (other ? 12.0 : 0.0)));
  auto go = [&](auto rdd) {
    using R = decltype(rdd);
    if (L.dual) {
      if (L.fine) {
        if (fam == 0) launch_one<true, 0, true, R>(S, rdd, count, maxlen, s.stream, variant, L, L.color_start[c]);
        else launch_one<true, 1, true, R>(S, rdd, count, maxlen, s.stream, variant, L, L.color_start[c]);
      } else {
        if (fam == 0) launch_one<false, 0, true, R>(S, rdd, count, maxlen, s.stream, variant, L, L.color_start[c]);
        else launch_one<false, 1, true, R>(S, rdd, count, maxlen, s.stream, variant, L, L.color_start[c]);
      }
    } else {
      if (L.fine) {
        if (fam == 0) launch_one<true, 0, false, R>(S, rdd, count, maxlen, s.stream, variant, L, L.color_start[c]);
        else launch_one<true, 1, false, R>(S, rdd, count, maxlen, s.stream, variant, L, L.color_start[c]);
      } else {
        if (fam == 0) launch_one<false, 0, false, R>(S, rdd, count, maxlen, s.stream, variant, L, L.color_start[c]);
        else launch_one<false, 1, false, R>(S, rdd, count, maxlen, s.stream, variant, L, L.color_start[c]);
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
// a rows buffer, and the staged ones whose LDS fits. With two colours only the split kernels: they
// gather every line's r before any line of the pass is solved, so lines of one colour that couple
// (two apart) read the lagged iterate, where a fused kernel would race.
std::vector<int> variants_for(bool fine, int len, bool has_rows, int ncol) {
  if (ncol == 2) {
    if (!has_rows) return {};
    return {kSplitVariant, kSplit4Variant, kFactorVariant};
  }
  std::vector<int> v{0};
  if (has_rows) {
    v.push_back(kSplitVariant);
    v.push_back(kSplit4Variant);
    v.push_back(kFactorVariant);
  }
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
  return variants_for(L.fine, maxlen, L.rows != nullptr, L.ncol);
}

// The rows buffer of the split variant: 18 planes of (most lines of one colour) x (longest line).
void smooth_alloc_rows(Solver& s, Level& L) {
  (void)s;
  if (L.rows) BD_HIP(hipFree(L.rows));
  if (L.du) BD_HIP(hipFree(L.du));
  if (L.fac) BD_HIP(hipFree(L.fac));
  L.rows = nullptr;
  L.du = nullptr;
  L.fac = nullptr;
  int most = 0;
  for (int c = 0; c < 2 * L.ncol; c++) most = std::max(most, L.color_start[c + 1] - L.color_start[c]);
  L.rows_cap = (size_t)most * ((L.max_len + 3) & ~3);  // room for either split variant's line stride
  BD_HIP(hipMalloc(&L.rows, 18 * L.rows_cap * sizeof(float)));
  L.du_stride = (L.max_len + 3) & ~3;
  L.du_cap = (size_t)L.color_start[2 * L.ncol] * L.du_stride;
  BD_HIP(hipMalloc(&L.du, 15 * L.du_cap * sizeof(float)));
  BD_HIP(hipMalloc(&L.fac, (size_t)kFacPlanes * L.du_cap * sizeof(float)));
}

// The line blocks of every line of level k (k_line_du per family), once its operator is final.
void smooth_line_blocks(Solver& s, int k) {
  Level& L = s.lv[k];
  if (!L.du) return;
  for (int fam = 1; fam >= 0; fam--) {
    // v-line colours come first in the table
    const int c0 = fam == 1 ? 0 : L.ncol, c1 = c0 + L.ncol;
    const int count = L.color_start[c1] - L.color_start[c0];
    if (count == 0) continue;
    int maxlen = 0;
    for (int l = 0; l < L.geo.nl; l++) maxlen = std::max(maxlen, fam == 1 ? L.geo.ny[l] : L.geo.nx[l]);
    const bool colmajor = L.dual && fam == 1;
    SmoothArgs S{colmajor ? L.opT : L.op, colmajor ? L.fixedT : L.fixed, L.Np, L.lines + L.color_start[c0],
                 colmajor ? L.xT : L.x, colmajor ? L.bT : L.b, nullptr, nullptr, L.geo, colmajor ? L.xT : L.x,
                 L.ncol == 2 ? 1 : 0};
    const dim3 grid((maxlen + 255) / 256, count);
    const int line0 = L.color_start[c0];
    auto go = [&](auto rdd) {
      using R = decltype(rdd);
      if (L.fine) {
        if (fam == 0) {
          if (L.dual) k_line_du<true, 0, true, R><<<grid, 256, 0, s.stream>>>(S, rdd, L.du, L.du_cap, L.du_stride, line0);
          else k_line_du<true, 0, false, R><<<grid, 256, 0, s.stream>>>(S, rdd, L.du, L.du_cap, L.du_stride, line0);
        } else {
          if (L.dual) k_line_du<true, 1, true, R><<<grid, 256, 0, s.stream>>>(S, rdd, L.du, L.du_cap, L.du_stride, line0);
          else k_line_du<true, 1, false, R><<<grid, 256, 0, s.stream>>>(S, rdd, L.du, L.du_cap, L.du_stride, line0);
        }
      } else {
        if (fam == 0) {
          if (L.dual) k_line_du<false, 0, true, R><<<grid, 256, 0, s.stream>>>(S, rdd, L.du, L.du_cap, L.du_stride, line0);
          else k_line_du<false, 0, false, R><<<grid, 256, 0, s.stream>>>(S, rdd, L.du, L.du_cap, L.du_stride, line0);
        } else {
          if (L.dual) k_line_du<false, 1, true, R><<<grid, 256, 0, s.stream>>>(S, rdd, L.du, L.du_cap, L.du_stride, line0);
          else k_line_du<false, 1, false, R><<<grid, 256, 0, s.stream>>>(S, rdd, L.du, L.du_cap, L.du_stride, line0);
        }
      }
    };
    if (L.oph)
      go(OpF16Dev{colmajor ? L.ophT : L.oph, L.Np, L.hs});
    else
      go(OpF32{colmajor ? L.opT : L.op, L.Np});
    // the family's factorizations from its line blocks
    const int block = ((maxlen + 3) / 4 + 63) / 64 * 64;
    const size_t shm = (size_t)24 * block * sizeof(float);
    if (fam == 0) {
      if (L.dual) k_line_factor<0, true><<<count, block, shm, s.stream>>>(S, L.du, L.fac, L.du_cap, L.du_stride, line0);
      else k_line_factor<0, false><<<count, block, shm, s.stream>>>(S, L.du, L.fac, L.du_cap, L.du_stride, line0);
    } else {
      if (L.dual) k_line_factor<1, true><<<count, block, shm, s.stream>>>(S, L.du, L.fac, L.du_cap, L.du_stride, line0);
      else k_line_factor<1, false><<<count, block, shm, s.stream>>>(S, L.du, L.fac, L.du_cap, L.du_stride, line0);
    }
  }
}

// Smooth all colours of level k (v-line colours first, then u-line colours; reversed for the
// post-smoothing sweep so the V-cycle stays symmetric).
void launch_smooth(Solver& s, int k, int reverse) {
  Level& L = s.lv[k];
  const int nc = 2 * L.ncol;
  for (int q = 0; q < nc; q++) {
    int c = reverse ? nc - 1 - q : q;
    const int fam = c < L.ncol ? 1 : 0;
    launch_colour(s, k, c, L.variant[fam], q == 1);
    // batch_copy: after a family's last pass, its layout's x to the other layout (on a one-colour level
    // the sweep's first pass wrote the pair's second buffers)
    const bool last = q == nc - 1 || (((reverse ? nc - 2 - q : q + 1) < L.ncol ? 1 : 0) != fam);
    if (L.batch_copy && last) {
      const bool colmajor = fam == 1;
      const bool pair = L.ncol == 1 && q == 0;
      const float* own = colmajor ? (pair ? L.xT2 : L.xT) : (pair ? L.x2 : L.x);
      float* other = colmajor ? (pair ? L.x2 : L.x) : (pair ? L.xT2 : L.xT);
      launch_layout_copy(s, k, own, other, !colmajor);
    }
  }
}

}  // namespace bd

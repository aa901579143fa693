// Exact solve of the coarsest multigrid level (BD_COARSE_EXACT). The reference smooths its coarsest
// level with coarse_sweeps colour-ordered line sweeps (newtonMultigrid.ts:495-528, gpu/solver.ts
// encodeVCycle); this option replaces them by A_c^-1 in FP64, the mixed-precision coarse solve.
//
// The coarsest level's DOF matrix (DOF 3 n + c, nodes row-major per layer) is banded with half
// bandwidth bw = 3 (2 nx + 2) + 2 (867 DOFs and bw = 110 for 17 x 17). Per hierarchy build:
//   k_band_load      the lower band from the symmetric coarse blocks (FP32 -> FP64);
//   k_band_cholesky_panel  right-looking banded Cholesky in one workgroup by panels of 32 columns
//                    in LDS (the diagonal block by one wave, the rows below one thread each) with
//                    rank-32 trailing updates in global memory;
//   k_band_inverse_blk  16 columns per workgroup, 16 rows per step (bands up to 240);
//   k_band_inverse   one wave per column c: L y = e_c, then L^T x = y, written as row c of the
//                    symmetric inverse (coalesced), for wider bands;
// and per V-cycle k_coarse_apply: one wave per row of A_c^-1 (FP64 dot with b, 6 MB at 867 DOFs).
#include <stdexcept>

#include "bd_internal.h"

namespace bd {

namespace {

__device__ __forceinline__ int sym_index(int r, int c) {
  // [xx, xy, xz, yy, yz, zz]
  const int lo = r < c ? r : c, hi = r < c ? c : r;
  return lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
}

struct BandArgs {
  const float* op;
  int N, Np;
  Geo geo;
  int n, bw;
  double* L;    // n x (bw + 1): L(i, j) at i * (bw + 1) + (i - j)
  double* inv;  // n x n
};

// lower band of the DOF matrix from the coarse blocks (centre + 12 forward blocks per row; a
// backward block is the transposed forward block of the other row, and the blocks are symmetric)
__global__ __launch_bounds__(256) void k_band_load(BandArgs B) {
  const int w = B.bw + 1;
  const size_t total = (size_t)B.n * w;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / w), o = (int)(e - (size_t)i * w);
    const int j = i - o;
    double v = 0.0;
    if (j >= 0) {
      const int a = i / 3, r = i - 3 * a, b = j / 3, c = j - 3 * b;
      const int la = geo_layer(B.geo, a), lb = geo_layer(B.geo, b);
      if (la == lb) {
        const int nx = B.geo.nx[la];
        const int ra = a - B.geo.off[la], rb = b - B.geo.off[la];
        const int ja = ra / nx, ia = ra - ja * nx, jb = rb / nx, ib = rb - jb * nx;
        const int di = ia - ib, dj = ja - jb;  // a relative to b (a follows b: dj > 0 or dj == 0, di >= 0)
        if (a == b) {
          v = B.op[sym_index(r, c) * B.Np + a];
        } else {
          const int k = coarse_block(di, dj);
          if (k > 0) v = B.op[(6 * k + sym_index(r, c)) * B.Np + b];
        }
      }
    }
    B.L[e] = v;
  }
}

// Right-looking banded Cholesky in one workgroup, blocked by panels of kPanel columns (a
// column-at-a-time form with the (bw + 1)^2 window in LDS measured 4.2 us per column at bw = 110,
// against 1.7 us for this one at bw = 207: it waits on a barrier and a global row load per column): the panel (its kPanel columns over the kPanel + bw rows they reach) is loaded to LDS;
// wave 0 alone factors its kPanel x kPanel diagonal block (a row per lane in registers, columns
// passed by scalar lane reads, no workgroup barriers); every row below then solves against that factor in registers, one thread per row; the
// panel is written back and the trailing band triangle (bw x bw) below it takes one rank-kPanel
// update in global memory (L2-resident), each thread holding a 4 x 4 tile of it in registers. Four
// workgroup barriers per panel.
constexpr int kPanel = 32;
// LDS row stride of the panel: odd in doubles, so that a column's rows fall in different banks
constexpr int kPanelStride = kPanel + 1;
// threads of the panel factorization (register room for a 32-wide row and a 4 x 4 tile pair)
constexpr int kPanelThreads = 512;

// lane `src`'s value of v (src uniform), for every lane
__device__ __forceinline__ double lane_read(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

__global__ __launch_bounds__(kPanelThreads) void k_band_cholesky_panel(BandArgs B) {
  extern __shared__ double pan[];  // (kPanel + bw) x kPanelStride, row-major
  const int n = B.n, bw = B.bw, w = bw + 1;
  const int t = threadIdx.x, T = blockDim.x;
  double* L = B.L;
  for (int k0 = 0; k0 < n; k0 += kPanel) {
    const int kb = min(kPanel, n - k0);
    const int rows = min(kb + bw, n - k0);  // panel rows k0 .. k0 + rows - 1
    // load: pan[r][c] = L(k0 + r, k0 + c) for c <= r within the band
    for (int e = t; e < rows * kPanel; e += T) {
      const int r = e / kPanel, c = e - r * kPanel;
      const int i = k0 + r, j = k0 + c;
      pan[r * kPanelStride + c] = (c < kb && j <= i && i - j <= bw) ? L[(size_t)i * w + (i - j)] : 0.0;
    }
    __syncthreads();
    // the diagonal block, right-looking, by wave 0: lane r holds row r in registers and column c's
    // values come from the lanes that own them (scalar lane reads)
    if (t < 64) {
      double a[kPanel];
#pragma unroll
      for (int c = 0; c < kPanel; c++) a[c] = (t < kb && c <= t) ? pan[t * kPanelStride + c] : 0.0;
#pragma unroll
      for (int c = 0; c < kPanel; c++) {
        if (c < kb) {
          const double d = lane_read(a[c], c);
          const double piv = d > 0.0 ? sqrt(d) : 1e-150;
          if (t == c) a[c] = piv;
          else if (t > c) a[c] /= piv;
#pragma unroll
          for (int c2 = c + 1; c2 < kPanel; c2++) {
            const double l = lane_read(a[c], c2);  // l(c2, c)
            if (t >= c2) a[c2] -= a[c] * l;
          }
        }
      }
      if (t < kb) {
#pragma unroll
        for (int c = 0; c < kPanel; c++)
          if (c <= t) pan[t * kPanelStride + c] = a[c];
      }
    }
    __syncthreads();
    // rows below the block: l_rc = (a_rc - sum_{c' < c} l_rc' l_cc') / l_cc, one thread per row
    for (int r = kb + t; r < rows; r += T) {
      double a[kPanel];
#pragma unroll
      for (int c = 0; c < kPanel; c++) a[c] = pan[r * kPanelStride + c];
#pragma unroll
      for (int c = 0; c < kPanel; c++) {
        if (c < kb) {
          double v = a[c];
#pragma unroll
          for (int c2 = 0; c2 < c; c2++) v -= a[c2] * pan[c * kPanelStride + c2];
          a[c] = v / pan[c * kPanelStride + c];
        }
      }
#pragma unroll
      for (int c = 0; c < kPanel; c++) pan[r * kPanelStride + c] = a[c];
    }
    __syncthreads();
    // write the panel back
    for (int e = t; e < rows * kPanel; e += T) {
      const int r = e / kPanel, c = e - r * kPanel;
      const int i = k0 + r, j = k0 + c;
      if (c < kb && j <= i && i - j <= bw) L[(size_t)i * w + (i - j)] = pan[r * kPanelStride + c];
    }
    // trailing update: L(i, j) -= sum_c pan[i - k0][c] pan[j - k0][c] for k0 + kb <= j <= i < k0 + rows
    const int m = rows - kb;  // trailing rows/columns
    const int tiles = (m + 3) / 4;
    const int ntile = tiles * (tiles + 1) / 2;
    for (int q = t; q < ntile; q += T) {
      int a = (int)((sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
      while ((a + 1) * (a + 2) / 2 <= q) a++;
      while (a * (a + 1) / 2 > q) a--;
      const int b = q - a * (a + 1) / 2;  // tile (a, b), b <= a
      // the tile's current values load while the rank-kPanel product is formed
      double acc[4][4], old[4][4];
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int i = k0 + kb + 4 * a + u, j = k0 + kb + 4 * b + v;
          old[u][v] = (i < k0 + rows && j <= i && i - j <= bw) ? L[(size_t)i * w + (i - j)] : 0.0;
          acc[u][v] = 0.0;
        }
      for (int c = 0; c < kb; c++) {
        double pi[4], pj[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int ri = kb + 4 * a + u, rj = kb + 4 * b + u;
          pi[u] = ri < rows ? pan[ri * kPanelStride + c] : 0.0;
          pj[u] = rj < rows ? pan[rj * kPanelStride + c] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
          for (int v = 0; v < 4; v++) acc[u][v] += pi[u] * pj[v];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int i = k0 + kb + 4 * a + u, j = k0 + kb + 4 * b + v;
          if (i < k0 + rows && j <= i && i - j <= bw) L[(size_t)i * w + (i - j)] = old[u][v] - acc[u][v];
        }
    }
    __syncthreads();
  }
}

// one wave per column c of the inverse: L y = e_c, L^T x = y; x is row c of the symmetric inverse.
// Each lane holds up to kInvPer band entries of the current row (bandwidth < 64 kInvPer) and the next
// row's are loaded before the current row's reduction; the last bandwidth values of y (forward) and
// x (backward) sit in a 256-entry LDS ring, y itself goes to the output row (read back, prefetched,
// in the backward pass). 4 KB of LDS per wave keeps every column's wave resident at once.
constexpr int kInvPer = 4;
constexpr int kRing = 64 * kInvPer;

__global__ __launch_bounds__(256) void k_band_recip(const double* L, int n, int w, double* rd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rd[i] = 1.0 / L[(size_t)i * w];
}

__global__ __launch_bounds__(64) void k_band_inverse(BandArgs B, const double* rd) {
  __shared__ double ring[kRing];
  const int n = B.n, bw = B.bw, w = bw + 1;
  const int c = blockIdx.x, lane = threadIdx.x;
  const double* L = B.L;
  double* out = B.inv + (size_t)c * n;
  for (int q = lane; q < kRing; q += 64) ring[q] = 0.0;
  for (int i = lane; i < c; i += 64) out[i] = 0.0;
  __syncthreads();
  double cur[kInvPer], nxt[kInvPer];
  // forward: row i uses L(i, k), k in [max(c, i - bw), i)
  auto load_fwd = [&](int i, double (&dst)[kInvPer]) {
    const int k0 = max(c, i - bw);
#pragma unroll
    for (int m = 0; m < kInvPer; m++) {
      const int k = k0 + lane + 64 * m;
      dst[m] = (i < n && k < i) ? L[(size_t)i * w + (i - k)] : 0.0;
    }
  };
  load_fwd(c, cur);
  double r_cur = rd[c], r_nxt = 0.0;
  for (int i = c; i < n; i++) {
    load_fwd(i + 1, nxt);
    if (i + 1 < n) r_nxt = rd[i + 1];
    const int k0 = max(c, i - bw);
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < kInvPer; m++) {
      const int k = k0 + lane + 64 * m;
      if (k < i) s += cur[m] * ring[k & (kRing - 1)];
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) {
      const double yi = ((i == c ? 1.0 : 0.0) - s) * r_cur;
      ring[i & (kRing - 1)] = yi;
      out[i] = yi;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kInvPer; m++) cur[m] = nxt[m];
    r_cur = r_nxt;
  }
  for (int q = lane; q < kRing; q += 64) ring[q] = 0.0;
  __syncthreads();
  // backward: row i uses L(k, i), k in (i, min(n - 1, i + bw)]; y_i is read back from the output row
  // (written by lane 0, which also reads it)
  auto load_bwd = [&](int i, double (&dst)[kInvPer]) {
    const int k1 = min(n - 1, i + bw);
#pragma unroll
    for (int m = 0; m < kInvPer; m++) {
      const int k = i + 1 + lane + 64 * m;
      dst[m] = (i >= 0 && k <= k1) ? L[(size_t)k * w + (k - i)] : 0.0;
    }
  };
  load_bwd(n - 1, cur);
  r_cur = rd[n - 1];
  double y_cur = lane == 0 ? out[n - 1] : 0.0, y_nxt = 0.0;
  for (int i = n - 1; i >= 0; i--) {
    load_bwd(i - 1, nxt);
    if (i > 0) {
      r_nxt = rd[i - 1];
      if (lane == 0) y_nxt = out[i - 1];
    }
    const int k1 = min(n - 1, i + bw);
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < kInvPer; m++) {
      const int k = i + 1 + lane + 64 * m;
      if (k <= k1) s += cur[m] * ring[k & (kRing - 1)];
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) {
      const double xi = (y_cur - s) * r_cur;
      ring[i & (kRing - 1)] = xi;
      out[i] = xi;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kInvPer; m++) cur[m] = nxt[m];
    r_cur = r_nxt;
    y_cur = y_nxt;
  }
}

// Blocked inverse: a workgroup of 256 threads solves kInvCols = 16 columns c of the inverse at once,
// kInvRows = 16 rows per step, so that each band row is read once per 16 columns and the serial
// chain is n / 16 steps per pass instead of n. Thread (cc, r) = (t / 16, t % 16) owns row i0 + r of
// column c0 + cc. A step: the band rows of the block (prefetched into registers during the previous
// step) go to LDS; each thread sums its row's band against the solution rows already known (an LDS
// ring of ring_mask + 1 >= bw + 16 rows x 16 columns); the 16 x 16 diagonal block is then solved by
// broadcasts within the column's 16 lanes. Forward L y = e_c from row c0 (y is zero above c), y to
// the output row; backward L^T x = y over all rows, y read back by the thread that wrote it.
constexpr int kInvCols = 16;
constexpr int kInvRows = 16;

// value of lane 16 g + q for every lane of 16-lane group g (q uniform): four scalar lane reads
// instead of an LDS permute round trip
__device__ __forceinline__ double group16_bcast(double v, int q) {
  const int g = (threadIdx.x & 63) >> 4;
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int l0 = __builtin_amdgcn_readlane(lo, q), l1 = __builtin_amdgcn_readlane(lo, 16 + q);
  const int l2 = __builtin_amdgcn_readlane(lo, 32 + q), l3 = __builtin_amdgcn_readlane(lo, 48 + q);
  const int h0 = __builtin_amdgcn_readlane(hi, q), h1 = __builtin_amdgcn_readlane(hi, 16 + q);
  const int h2 = __builtin_amdgcn_readlane(hi, 32 + q), h3 = __builtin_amdgcn_readlane(hi, 48 + q);
  const int rl = g == 0 ? l0 : g == 1 ? l1 : g == 2 ? l2 : l3;
  const int rh = g == 0 ? h0 : g == 1 ? h1 : g == 2 ? h2 : h3;
  return __hiloint2double(rh, rl);
}

__global__ __launch_bounds__(256) void k_band_inverse_blk(BandArgs B, const double* rd, int ring_mask) {
  extern __shared__ double sh[];
  const int n = B.n, bw = B.bw, w = bw + 1, lw = bw + kInvRows;
  const int ring_rows = ring_mask + 1;
  double* ring = sh;                        // ring_rows x kInvCols
  double* Lb = sh + ring_rows * kInvCols;   // kInvRows x lw (forward) or lw x kInvRows (backward)
  const double* L = B.L;
  const int c0 = blockIdx.x * kInvCols;
  const int t = threadIdx.x, cc = t >> 4, r = t & 15;
  const int c = c0 + cc;
  double* out = B.inv + (size_t)(c < n ? c : 0) * n;
  const bool col = c < n;
  constexpr int kPer = 16;  // band values a thread stages per step (>= kInvRows * lw / 256)
  double pre[kPer];

  for (int q = t; q < ring_rows * kInvCols; q += 256) ring[q] = 0.0;
  if (col)
    for (int i = r; i < c0; i += kInvRows) out[i] = 0.0;

  // forward: Lb[rr * lw + j] = L(i0 + rr, i0 - bw + j)
  auto fetch_fwd = [&](int i0) {
#pragma unroll
    for (int m = 0; m < kPer; m++) {
      const int q = t + 256 * m;
      const int rr = q / lw, j = q - rr * lw;
      const int i = i0 + rr, k = i0 - bw + j, d = i - k;
      pre[m] = (q < kInvRows * lw && i < n && k >= 0 && d >= 0 && d <= bw) ? L[(size_t)i * w + d] : 0.0;
    }
  };
  fetch_fwd(c0);
  double rd_next = c0 + r < n ? rd[c0 + r] : 0.0;
  for (int i0 = c0; i0 < n; i0 += kInvRows) {
    __syncthreads();  // the previous step's Lb and ring reads are done
#pragma unroll
    for (int m = 0; m < kPer; m++) {
      const int q = t + 256 * m;
      if (q < kInvRows * lw) Lb[q] = pre[m];
    }
    __syncthreads();
    if (i0 + kInvRows < n) fetch_fwd(i0 + kInvRows);
    const int i = i0 + r;
    const double rdi = rd_next;
    rd_next = i + kInvRows < n ? rd[i + kInvRows] : 0.0;
    const double* Lr = Lb + r * lw;
    // four independent sums keep the LDS reads of the 200-odd terms in flight
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int j = r;
    for (; j + 3 < bw; j += 4) {
      s0 += Lr[j] * ring[((i0 - bw + j) & ring_mask) * kInvCols + cc];
      s1 += Lr[j + 1] * ring[((i0 - bw + j + 1) & ring_mask) * kInvCols + cc];
      s2 += Lr[j + 2] * ring[((i0 - bw + j + 2) & ring_mask) * kInvCols + cc];
      s3 += Lr[j + 3] * ring[((i0 - bw + j + 3) & ring_mask) * kInvCols + cc];
    }
    for (; j < bw; j++) s0 += Lr[j] * ring[((i0 - bw + j) & ring_mask) * kInvCols + cc];
    const double s = (s0 + s1) + (s2 + s3);
    double v = (i == c ? 1.0 : 0.0) - s, y = 0.0;
#pragma unroll
    for (int q = 0; q < kInvRows; q++) {
      const double yq = group16_bcast(v * rdi, q);
      if (r == q) y = yq;
      if (r > q) v -= Lr[bw + q] * yq;
    }
    ring[(i & ring_mask) * kInvCols + cc] = y;
    if (col && i < n) out[i] = y;
  }
  __syncthreads();
  for (int q = t; q < ring_rows * kInvCols; q += 256) ring[q] = 0.0;

  // backward: Lb[kk * kInvRows + q] = L(i0 + kk, i0 + q) for 0 < kk - q <= bw
  auto fetch_bwd = [&](int i0) {
#pragma unroll
    for (int m = 0; m < kPer; m++) {
      const int e = t + 256 * m;
      const int kk = e / kInvRows, q = e - kk * kInvRows;
      const int k = i0 + kk, d = kk - q;
      pre[m] = (e < lw * kInvRows && k < n && d > 0 && d <= bw) ? L[(size_t)k * w + d] : 0.0;
    }
  };
  const int last = ((n - 1) / kInvRows) * kInvRows;
  fetch_bwd(last);
  rd_next = last + r < n ? rd[last + r] : 0.0;
  double y_next = (col && last + r < n) ? out[last + r] : 0.0;
  for (int i0 = last; i0 >= 0; i0 -= kInvRows) {
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kPer; m++) {
      const int e = t + 256 * m;
      if (e < lw * kInvRows) Lb[e] = pre[m];
    }
    __syncthreads();
    if (i0 > 0) fetch_bwd(i0 - kInvRows);
    const int i = i0 + r;
    const double rdi = rd_next, yi = y_next;
    if (i0 > 0) {
      rd_next = rd[i - kInvRows];
      y_next = col ? out[i - kInvRows] : 0.0;
    }
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int kk = kInvRows;
    for (; kk + 3 < lw; kk += 4) {
      s0 += Lb[kk * kInvRows + r] * ring[((i0 + kk) & ring_mask) * kInvCols + cc];
      s1 += Lb[(kk + 1) * kInvRows + r] * ring[((i0 + kk + 1) & ring_mask) * kInvCols + cc];
      s2 += Lb[(kk + 2) * kInvRows + r] * ring[((i0 + kk + 2) & ring_mask) * kInvCols + cc];
      s3 += Lb[(kk + 3) * kInvRows + r] * ring[((i0 + kk + 3) & ring_mask) * kInvCols + cc];
    }
    for (; kk < lw; kk++) s0 += Lb[kk * kInvRows + r] * ring[((i0 + kk) & ring_mask) * kInvCols + cc];
    const double s = (s0 + s1) + (s2 + s3);
    double v = yi - s, x = 0.0;
#pragma unroll
    for (int q = kInvRows - 1; q >= 0; q--) {
      const double xq = group16_bcast(v * rdi, q);
      if (r == q) x = xq;
      if (r < q) v -= Lb[q * kInvRows + r] * xq;
    }
    ring[(i & ring_mask) * kInvCols + cc] = x;
    if (col && i < n) out[i] = x;
  }
}

// x = A_c^-1 b on the coarsest level: one wave per row, FP64 products and sums
__global__ __launch_bounds__(256) void k_coarse_apply(const double* inv, int n, int Np, const float* b, float* x,
                                                      const int* gate) {
  if (gate && *gate) return;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n) return;
  const double* r = inv + (size_t)row * n;
  double s = 0.0;
  for (int j = lane; j < n; j += 64) {
    const int node = j / 3, c = j - 3 * node;
    s += r[j] * (double)b[c * Np + node];
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) {
    const int node = row / 3, c = row - 3 * node;
    x[c * Np + node] = (float)s;
  }
}

int band_width(const Geo& g) {
  int bw = 2;
  for (int l = 0; l < g.nl; l++) bw = std::max(bw, 3 * (2 * g.nx[l] + 2) + 2);
  return bw;
}

}  // namespace

size_t coarse_exact_bytes(const Level& L) {
  const size_t n = 3 * (size_t)L.N, w = band_width(L.geo) + 1;
  return sizeof(double) * (n * w + n * n);
}

void coarse_exact_alloc(Solver& s) {
  Level& L = s.lv[s.nlev - 1];
  const int n = 3 * L.N, bw = band_width(L.geo);
  if (bw >= kRing) throw std::runtime_error("exact coarse solve: coarsest band wider than 255");
  BD_HIP(hipMalloc(&s.cband, sizeof(double) * (size_t)n * (bw + 1)));
  BD_HIP(hipMalloc(&s.cinv, sizeof(double) * (size_t)n * n));
  BD_HIP(hipMalloc(&s.crd, sizeof(double) * (size_t)n));
  const size_t plds = sizeof(double) * (size_t)(kPanel + bw) * kPanelStride;
  if (plds > 150 * 1024) throw std::runtime_error("exact coarse solve: coarsest band too wide");
  BD_HIP(hipFuncSetAttribute((const void*)k_band_cholesky_panel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)plds));
  s.cn = n;
  s.cbw = bw;
}

void launch_coarse_factor(Solver& s) {
  Level& L = s.lv[s.nlev - 1];
  BandArgs B{L.op, L.N, L.Np, L.geo, s.cn, s.cbw, s.cband, s.cinv};
  const int w = s.cbw + 1;
  {
    // band load + factor (the band in and out once)
    ProfScope ps(s, PC_COARSE_FACTOR, 8.0 * ((double)s.cn * w * 3));
    k_band_load<<<grid_for(s.cn * w, 256), 256, 0, s.stream>>>(B);
    k_band_cholesky_panel<<<1, kPanelThreads, sizeof(double) * (size_t)(kPanel + s.cbw) * kPanelStride, s.stream>>>(B);
  }
  {
    // inverse: the band read per column, the inverse written once
    ProfScope ps(s, PC_COARSE_INVERSE, 8.0 * (double)s.cn * s.cn);
    k_band_recip<<<grid_for(s.cn, 256), 256, 0, s.stream>>>(s.cband, s.cn, w, s.crd);
    if (s.cbw + kInvRows <= 256) {
      // the blocked kernel's ring (256 rows) and staging (16 values per thread) cover the band
      const size_t shm = sizeof(double) * (256 * kInvCols + (size_t)kInvRows * (s.cbw + kInvRows));
      k_band_inverse_blk<<<(s.cn + kInvCols - 1) / kInvCols, 256, shm, s.stream>>>(B, s.crd, 255);
    } else {
      k_band_inverse<<<s.cn, 64, 0, s.stream>>>(B, s.crd);
    }
  }
}

void launch_coarse_apply(Solver& s) {
  Level& L = s.lv[s.nlev - 1];
  ProfScope ps(s, PC_COARSE_APPLY, 8.0 * (double)s.cn * s.cn + 24.0 * L.N);
  k_coarse_apply<<<(s.cn + 3) / 4, 256, 0, s.stream>>>(s.cinv, s.cn, L.Np, L.b, L.x, &s.scal->done);
}

}  // namespace bd

// Galerkin multigrid on the half-stencil operators: connectivity-masked prolongation weights,
// the Galerkin triple product P^T A P, residual / restriction / prolongation and the V-cycle
// (the line smoother is in smooth.hip).
//
// Restates buildProlongation / galerkin / vcycle (src/simulation/reference/newtonMultigrid.ts:
// 287-387, 495-528) and bracket / coarseToFine (lattice.ts:13-41).
#include <stdexcept>

#include "bd_internal.h"
#include "mg_dev.h"
#include "oprd.h"
#include "reduce.h"

namespace bd {

namespace {

template <bool FINE, class R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_residual(LvArgs a, R rd, const float* x,
                                                                                              const float* b, float* r,
                                                                                              const int* gate) {
  if (gate && *gate) return;
  int n = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= a.N) return;
  const int Np = a.Np;
  if (a.fixed[n]) {
    r[n] = r[Np + n] = r[2 * Np + n] = 0.f;
    return;
  }
  float y[3];
  vc_apply_row<FINE>(rd.live(), a.geo, x, Np, n, y);
  r[n] = b[n] - y[0];
  r[Np + n] = b[Np + n] - y[1];
  r[2 * Np + n] = b[2 * Np + n] - y[2];
}

// The FP16 scale of a level: the largest diagonal entry of a free row (it bounds every entry of the
// symmetric positive definite operator), as float bits through atomicMax (non-negative floats order
// like their bits).
__global__ __launch_bounds__(256) void k_half_max(const float* op, const uint8_t* fixed, int N, int Np, float* hs) {
  __shared__ float red[256];
  float m = 0.f;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x)
    if (!fixed[n]) m = fmaxf(m, fmaxf(fabsf(op[n]), fmaxf(fabsf(op[3 * Np + n]), fabsf(op[5 * Np + n]))));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(reinterpret_cast<unsigned*>(hs + 2), __float_as_uint(red[0]));
}

// scale 2^e with the largest free diagonal entry in (2^13, 2^14] (1 when the level has none)
__device__ __forceinline__ float half_scale(const float* hs) {
  const float mx = __uint_as_float(*reinterpret_cast<const unsigned*>(hs + 2));
  if (!(mx > 0.f)) return 1.f;
  int ex;
  const float fr = frexpf(mx, &ex);  // mx = fr 2^ex, fr in [0.5, 1)
  const int c = fr == 0.5f ? ex - 1 : ex;  // ceil(log2 mx)
  return ldexpf(1.f, 14 - c);
}

// planes 2w and 2w + 1 of a level operator -> word w (FP16, round to nearest even), times the scale
__global__ __launch_bounds__(256) void k_pack_half(const float* op, int words, int Np, float* hs, uint32_t* out) {
  const float sc = half_scale(hs);
  const size_t total = (size_t)words * Np;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    hs[0] = sc;
    hs[1] = 1.f / sc;
  }
  if (t >= total) return;
  const size_t w = t / Np, n = t - w * Np;
  // saturated at the largest finite half: a fixed row (outside the scale's maximum) never becomes inf,
  // so a masked product with it stays zero
  const float a = fminf(fmaxf(op[(2 * w) * Np + n] * sc, -65504.f), 65504.f);
  const float b = fminf(fmaxf(op[(2 * w + 1) * Np + n] * sc, -65504.f), 65504.f);
  const _Float16 lo = (_Float16)a, hi = (_Float16)b;
  out[t] = (uint32_t)__builtin_bit_cast(uint16_t, lo) | ((uint32_t)__builtin_bit_cast(uint16_t, hi) << 16);
}

// ---------------------------------------------------------------- prolongation (lattice.ts:13-41)
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
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_restrict(XferArgs X) {
  if (X.gate && *X.gate) return;
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= X.Nc) return;
  const int Npc = X.Npc;
  float s[3];
  restrict_node(X, I, s);
  for (int q = 0; q < 3; q++) {
    X.bc[q * Npc + I] = s[q];
    if (X.xc) X.xc[q * Npc + I] = 0.f;    // null for a plain restriction of a vector
    if (X.xcT) X.xcT[q * Npc + I] = 0.f;  // every entry of the permuted copy is cleared as well
  }
  if (X.bcT) {
    // b's column-major copy for the level's v-line passes, written here rather than by a transpose
    // launch per visit (a strided store per node; the coarse levels are small)
    int l, i, j;
    node_ij(X.cgeo, I, l, i, j);
    const int cm = X.cgeo.off[l] + i * X.cgeo.ny[l] + j;
    for (int q = 0; q < 3; q++) X.bcT[q * Npc + cm] = s[q];
  }
}

// levels without a column-major copy: one node per thread
__global__ __launch_bounds__(256) void k_prolong_add(XferArgs X) {
  if (X.gate && *X.gate) return;
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= X.a.N) return;
  if (X.a.fixed[f]) return;
  int l, i, j;
  node_ij(X.a.geo, f, l, i, j);
  float s[3];
  prolong_sum(X, f, l, i, j, s);
  for (int q = 0; q < 3; q++) X.xf[q * X.a.Np + f] += s[q];
}

// number of 32 x 32 node tiles of a level (every layer tiled on its own)
int prolong_tiles(const Geo& g) {
  int n = 0;
  for (int l = 0; l < g.nl; l++) n += ((g.nx[l] + 31) / 32) * ((g.ny[l] + 31) / 32);
  return n;
}

// levels with a column-major copy: one 32 x 32 node tile of one layer per block (thread (tx, ty)
// takes rows ty, ty + 8, ...). x is updated row by row; the copy is written from LDS along
// columns, 32 consecutive nodes (128 bytes) per column, where a node-per-thread store puts every
// lane of a wave in its own cache line. Fixed nodes write nothing to either copy.
__global__ __launch_bounds__(256) void k_prolong_add_tiled(XferArgs X) {
  if (X.gate && *X.gate) return;
  __shared__ float tile[3][32][33];
  __shared__ int live[32][33];
  int b = blockIdx.x, l = 0, tnx = 1;
  for (; l < X.a.geo.nl; l++) {
    tnx = (X.a.geo.nx[l] + 31) >> 5;
    const int nt = tnx * ((X.a.geo.ny[l] + 31) >> 5);
    if (b < nt) break;
    b -= nt;
  }
  if (l >= X.a.geo.nl) return;  // the grid holds exactly the level's tiles: not reached
  const int nx = X.a.geo.nx[l], ny = X.a.geo.ny[l], off = X.a.geo.off[l], Np = X.a.Np;
  const int i0 = (b % tnx) * 32, j0 = (b / tnx) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, j = j0 + r;
    int on = 0;
    float v[3] = {0.f, 0.f, 0.f};
    if (i < nx && j < ny) {
      const int f = off + j * nx + i;
      if (!X.a.fixed[f]) {
        float s[3];
        prolong_sum(X, f, l, i, j, s);
#pragma unroll
        for (int q = 0; q < 3; q++) {
          v[q] = X.xf[q * Np + f] + s[q];
          X.xf[q * Np + f] = v[q];
        }
        on = 1;
      }
    }
#pragma unroll
    for (int q = 0; q < 3; q++) tile[q][r][tx] = v[q];
    live[r][tx] = on;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int j = j0 + tx, i = i0 + r;  // column i, row j: tile entry [j - j0][i - i0]
    if (i < nx && j < ny && live[tx][r]) {
#pragma unroll
      for (int q = 0; q < 3; q++) X.xfT[q * Np + off + i * ny + j] = tile[q][tx][r];
    }
  }
}

__global__ void k_zero(float* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// row-major (off + j nx + i) -> column-major (off + i ny + j) copy of `planes` planes of one layer,
// through 32 x 32 tiles in LDS so both sides are coalesced
template <class T>
__global__ __launch_bounds__(256) void k_to_colmajor(const T* in, T* out, int Np, int off, int nx, int ny) {
  __shared__ T tile[32][33];
  const size_t q = blockIdx.z;
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const T* src = in + q * Np + off;
  T* dst = out + q * Np + off;
  for (int r = threadIdx.y; r < 32; r += 8) {
    const int i = i0 + threadIdx.x, j = j0 + r;
    if (i < nx && j < ny) tile[r][threadIdx.x] = src[j * nx + i];
  }
  __syncthreads();
  for (int r = threadIdx.y; r < 32; r += 8) {
    const int j = j0 + threadIdx.x, i = i0 + r;
    if (i < nx && j < ny) dst[i * ny + j] = tile[threadIdx.x][r];
  }
}

// column-major -> row-major copy (the inverse of k_to_colmajor), same tiling
template <class T>
__global__ __launch_bounds__(256) void k_to_rowmajor(const T* in, T* out, int Np, int off, int nx, int ny) {
  __shared__ T tile[32][33];
  const size_t q = blockIdx.z;
  const int j0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
  const T* src = in + q * Np + off;
  T* dst = out + q * Np + off;
  for (int r = threadIdx.y; r < 32; r += 8) {
    const int j = j0 + threadIdx.x, i = i0 + r;
    if (i < nx && j < ny) tile[r][threadIdx.x] = src[i * ny + j];
  }
  __syncthreads();
  for (int r = threadIdx.y; r < 32; r += 8) {
    const int i = i0 + threadIdx.x, j = j0 + r;
    if (i < nx && j < ny) dst[j * nx + i] = tile[threadIdx.x][r];
  }
}

template <class T>
void to_colmajor(Solver& s, const Level& L, const T* in, T* out, int planes) {
  for (int l = 0; l < L.geo.nl; l++) {
    const int nx = L.geo.nx[l], ny = L.geo.ny[l];
    dim3 grid((nx + 31) / 32, (ny + 31) / 32, planes);
    k_to_colmajor<T><<<grid, dim3(32, 8), 0, s.stream>>>(in, out, L.Np, L.geo.off[l], nx, ny);
  }
}

LvArgs lv_args(const Level& L) { return LvArgs{L.op, L.fixed, L.N, L.Np, L.geo}; }

const int* pcg_gate(Solver& s) { return &s.scal->done; }

}  // namespace

void launch_zero(Solver& s, float* p, size_t n) {
  if (n == 0) return;
  k_zero<<<(unsigned)((n + 255) / 256), 256, 0, s.stream>>>(p, n);
}

// Refreshes a level operator's derived copies once it is final (after assembly or RAP): the
// column-major copy of dual-layout levels and, under BD_PRECISION_FP16, the FP16 pairs of both.
void launch_colmajor_op(Solver& s, int k) {
  Level& L = s.lv[k];
  const int comps = L.fine ? kFineComps : kCoarseComps;
  if (L.dual) {
    ProfScope ps(s, PC_TRANSPOSE, 2.0 * L.N * (4.0 * comps + 1.0));
    to_colmajor<float>(s, L, L.op, L.opT, comps);
    to_colmajor<uint8_t>(s, L, L.fixed, L.fixedT, 1);
  }
  if (L.oph) {
    const int words = comps / 2;
    const size_t total = (size_t)words * L.Np;
    ProfScope ps(s, PC_TRANSPOSE, (double)total * (L.dual ? 24.0 : 12.0));
    launch_zero(s, L.hs, 4);
    k_half_max<<<std::min(1024, grid_for(L.N, 256)), 256, 0, s.stream>>>(L.op, L.fixed, L.N, L.Np, L.hs);
    k_pack_half<<<(unsigned)((total + 255) / 256), 256, 0, s.stream>>>(L.op, words, L.Np, L.hs, L.oph);
    if (L.dual) k_pack_half<<<(unsigned)((total + 255) / 256), 256, 0, s.stream>>>(L.opT, words, L.Np, L.hs, L.ophT);
  }
  if (L.du) smooth_line_blocks(s, k);
}

void launch_layout_copy(Solver& s, int k, const float* src, float* dst, bool to_col) {
  Level& L = s.lv[k];
  if (!L.dual) return;
  ProfScope ps(s, PC_TRANSPOSE, 2.0 * L.N * 12.0);
  if (to_col) {
    to_colmajor<float>(s, L, src, dst, 3);
    return;
  }
  for (int l = 0; l < L.geo.nl; l++) {
    const int nx = L.geo.nx[l], ny = L.geo.ny[l];
    dim3 grid((ny + 31) / 32, (nx + 31) / 32, 3);
    k_to_rowmajor<float><<<grid, dim3(32, 8), 0, s.stream>>>(src, dst, L.Np, L.geo.off[l], nx, ny);
  }
}

void launch_colmajor_x(Solver& s, int k) {
  Level& L = s.lv[k];
  if (!L.dual) return;
  to_colmajor<float>(s, L, L.x, L.xT, 3);
}

// b in the column-major layout, for the v-line passes of a dual-layout level (b is fixed while the
// level is smoothed: the V-cycle refreshes it once per visit, before the pre-smoothing)
void launch_colmajor_b(Solver& s, int k) {
  Level& L = s.lv[k];
  if (!L.dual) return;
  ProfScope ps(s, PC_TRANSPOSE, 2.0 * L.N * 12.0);
  to_colmajor<float>(s, L, L.b, L.bT, 3);
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
  // unique bytes: the node's forward planes (32 fine, 78 coarse) + fixed flag + x, b read and r written
  ProfScope ps(s, L.fine ? PC_RESIDUAL_FINE : PC_RESIDUAL_COARSE, L.N * (L.fine ? 165.0 : 349.0));
  const unsigned g = grid_for(L.N, 256);
  if (L.oph) {
    // the FP16 pairs; the level's inverse scale is read from device memory by the kernel
    const OpF16Dev rd{L.oph, L.Np, L.hs};
    if (L.fine)
      k_residual<true><<<g, 256, 0, s.stream>>>(lv_args(L), rd, x, b, r, pcg_gate(s));
    else
      k_residual<false><<<g, 256, 0, s.stream>>>(lv_args(L), rd, x, b, r, pcg_gate(s));
  } else {
    const OpF32 rd{L.op, L.Np};
    if (L.fine)
      k_residual<true><<<g, 256, 0, s.stream>>>(lv_args(L), rd, x, b, r, pcg_gate(s));
    else
      k_residual<false><<<g, 256, 0, s.stream>>>(lv_args(L), rd, x, b, r, pcg_gate(s));
  }
}

void launch_restrict(Solver& s, int k) {
  Level& L = s.lv[k];
  Level& Cl = s.lv[k + 1];
  XferArgs X{lv_args(L), Cl.geo, Cl.N, Cl.Np, L.pw, Cl.fixed, L.r, Cl.b, Cl.x, L.x, pcg_gate(s),
             Cl.dual ? Cl.xT : nullptr, nullptr, Cl.dual ? Cl.bT : nullptr};
  ProfScope ps(s, PC_RESTRICT, L.N * 29.0 + Cl.N * (Cl.dual ? 49.0 : 25.0));
  k_restrict<<<grid_for(Cl.N, 256), 256, 0, s.stream>>>(X);
}

void launch_prolong_add(Solver& s, int k) {
  Level& L = s.lv[k];
  Level& Cl = s.lv[k + 1];
  XferArgs X{lv_args(L), Cl.geo, Cl.N, Cl.Np, L.pw, Cl.fixed, L.r, Cl.b, Cl.x, L.x, pcg_gate(s),
             nullptr, L.dual ? L.xT : nullptr};
  ProfScope ps(s, PC_PROLONG, L.N * (L.dual ? 53.0 : 41.0) + Cl.N * 12.0);
  if (L.dual)
    k_prolong_add_tiled<<<prolong_tiles(L.geo), 256, 0, s.stream>>>(X);
  else
    k_prolong_add<<<grid_for(L.N, 256), 256, 0, s.stream>>>(X);
}

namespace {

// y = A x on a level (fixed rows 0) and x^T A x in FP64; the reduction's block partials, its ticket
// and the result live in the caller's scratch (kDotBlocks doubles, then the ticket, then the result)
template <bool FINE>
__global__ __launch_bounds__(1024) void k_apply_dot(LvArgs a, const float* x, float* y, double* scratch, double* out) {
  double v[1] = {0};
  const int Np = a.Np;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < a.N; n += gridDim.x * blockDim.x) {
    float r[3] = {0.f, 0.f, 0.f};
    if (!a.fixed[n]) apply_row<FINE>(a, x, n, r);
    for (int q = 0; q < 3; q++) {
      y[q * Np + n] = r[q];
      v[0] += (double)x[q * Np + n] * (double)r[q];
    }
  }
  double tot[1];
  if (grid_sum<1>(v, scratch, reinterpret_cast<unsigned int*>(scratch + kDotBlocks), tot)) *out = tot[0];
}

__global__ void k_clear_ticket(double* scratch) { *reinterpret_cast<unsigned int*>(scratch + kDotBlocks) = 0u; }

// y = A x on a level with every product and sum in FP64 (fixed rows 0): for smooth x the row sums
// cancel to a small fraction of their terms, which FP32 accumulation would not resolve
template <bool FINE>
__global__ __launch_bounds__(256) void k_apply_f64(LvArgs a, const float* x, float* y) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.N) return;
  const int Np = a.Np;
  double r[3] = {0, 0, 0};
  if (!a.fixed[n]) {
    int l, i, j;
    node_ij(a.geo, n, l, i, j);
    auto add = [&](const float* b, int m) {
      const double x0 = x[m], x1 = x[Np + m], x2 = x[2 * Np + m];
      r[0] += (double)b[0] * x0 + (double)b[1] * x1 + (double)b[2] * x2;
      r[1] += (double)b[1] * x0 + (double)b[3] * x1 + (double)b[4] * x2;
      r[2] += (double)b[2] * x0 + (double)b[4] * x1 + (double)b[5] * x2;
    };
    float c[6];
    blk_load(a.op, Np, 0, n, c);
    add(c, n);
    for_couplings<FINE>(a.op, Np, n, i, j, a.geo.nx[l], a.geo.ny[l], [&](int, int, int m, const float* b) { add(b, m); });
  }
  for (int q = 0; q < 3; q++) y[q * Np + n] = (float)r[q];
}

}  // namespace

void launch_apply_dot(Solver& s, int k, const float* x, float* y, double* scratch, double* xax, hipStream_t st) {
  Level& L = s.lv[k];
  const int g = std::min(kDotBlocks, grid_for(L.N, 1024));
  k_clear_ticket<<<1, 1, 0, st>>>(scratch);
  if (L.fine)
    k_apply_dot<true><<<g, 1024, 0, st>>>(lv_args(L), x, y, scratch, xax);
  else
    k_apply_dot<false><<<g, 1024, 0, st>>>(lv_args(L), x, y, scratch, xax);
}

void launch_apply_f64(Solver& s, int k, const float* x, float* y) {
  Level& L = s.lv[k];
  if (L.fine)
    k_apply_f64<true><<<grid_for(L.N, 256), 256, 0, s.stream>>>(lv_args(L), x, y);
  else
    k_apply_f64<false><<<grid_for(L.N, 256), 256, 0, s.stream>>>(lv_args(L), x, y);
}

void launch_restrict_vec(Solver& s, int k, const float* r, float* bc, const int* gate) {
  Level& L = s.lv[k];
  Level& Cl = s.lv[k + 1];
  XferArgs X{lv_args(L), Cl.geo, Cl.N, Cl.Np, L.pw, Cl.fixed, r, bc, nullptr, nullptr, gate, nullptr, nullptr};
  k_restrict<<<grid_for(Cl.N, 256), 256, 0, s.stream>>>(X);
}

void launch_prolong_vec(Solver& s, int k, const float* xc, float* xf, bool set, const int* gate) {
  Level& L = s.lv[k];
  Level& Cl = s.lv[k + 1];
  if (set) launch_zero(s, xf, 3 * (size_t)L.Np);
  XferArgs X{lv_args(L), Cl.geo, Cl.N, Cl.Np, L.pw, Cl.fixed, nullptr, nullptr, const_cast<float*>(xc), xf, gate, nullptr,
             nullptr};
  k_prolong_add<<<grid_for(L.N, 256), 256, 0, s.stream>>>(X);
}

// vcycle (newtonMultigrid.ts:495-528) on level k's x / b buffers, x zero on entry (below the fine
// level, the restriction that filled b also wrote its column-major copy)
void vcycle(Solver& s, int k) {
  Level& L = s.lv[k];
  if (k == s.nlev - 1) {
    if (coarse_exact(s)) {
      coarse_join(s);
      launch_coarse_apply(s);
      return;
    }
    if (k == 0) launch_colmajor_b(s, k);
    for (int sw = 0; sw < s.P.coarse_sweeps; sw++) launch_smooth(s, k, sw & 1);
    return;
  }
  if (k == 0) launch_colmajor_b(s, k);
  for (int sw = 0; sw < s.P.smooth_sweeps; sw++) launch_smooth(s, k, 0);
  launch_residual(s, k, L.x, L.b, L.r);
  launch_restrict(s, k);
  vcycle(s, k + 1);
  launch_prolong_add(s, k);
  for (int sw = 0; sw < s.P.smooth_sweeps; sw++) launch_smooth(s, k, 1);
}

}  // namespace bd

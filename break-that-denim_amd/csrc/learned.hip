// The learned coarse-space correction: a mesh message-passing network on one coarse level of the
// Galerkin hierarchy predicts kNetK displacement fields Z (3 planes each) from that level's
// operator. Z serves twice in the Newton solve:
//   warm start          d0 = P Z (Z^T A_l Z)^-1 Z^T R b        (A-optimal start in span(P Z))
//   coarse correction   z = V(r) + P Z (Z^T A_l Z)^-1 Z^T R r   (additive, inside flexible PCG)
// where R = P^T restricts the fine residual down to the net's level l and P prolongs back, and
// A_l = P^T A P is the level's Galerkin operator, so Z^T A_l Z is the Gram matrix of the prolonged
// fields. Inside PCG the fields deflate: with W = A P Z stored per Newton system and
// G = (P Z)^T W, every preconditioned residual is projected, z <- z - P Z G^-1 W^T z, which with
// the warm start keeps the residual orthogonal to span(P Z) (deflated CG: Saad, Yeung, Erhel and
// Guyomarc'h, SIAM J. Sci. Comput. 21, 2000); fields that capture little cannot make the iteration
// worse than without them, the ones that capture slow modes remove them. The network and its
// training are this build's own (the reference has no learned part).
//
// Network (BF16 weights and activations, FP32 accumulation on the matrix cores, MFMA
// 16x16x32 bf16): per level node, 32 input features (lattice position encodings, fixed flag, the
// operator's diagonal block, the restricted Newton right-hand side; 20 used, 12 zero) -> encoder 32 -> 64 (ReLU) -> 4 message-passing layers
// h += ReLU([h, mean of the 4 lattice neighbours' h] W + b) (128 -> 64) -> decoder 64 -> 32 (24 used:
// kNetK fields x 3 components).
//
// Training (on the GPU, from samples the plain solver records): per sample the level operator A_l
// and target right-hand sides b_t = R A e_t with s_t = e_t^T A e_t, for e_t the Newton direction
// (the warm-start target) and the error the plain solve still had after kSlowIter iterations (the
// slow modes the coarse correction must capture). The loss is the fraction of each target's
// energy that span(P Z) misses,
//   L = sum_t w_t (1 - c_t^T y_t / s_t),  c_t = Z^T b_t,  y_t = (Z^T A_l Z)^+ c_t,
// whose gradient with respect to Z is dL/dZ = -sum_t (2 w_t / s_t) (b_t - A_l Z y_t) y_t^T.
// Backward through the layers runs on the same matrix cores; Adam updates FP32 master weights.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <vector>

#include "bd_internal.h"
#include "reduce.h"

namespace bd {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 32, kHid = 64, kMp = 4, kOut = 32;
constexpr int kNodesPerBlock = 64;  // 4 waves x 16 nodes
constexpr int kTargets = 2;

// layer table of the flat parameter vector (weights in-major W[i][o], then the bias)
struct LayerDesc {
  int in, out, w, b;
};
constexpr LayerDesc kLayers[kMp + 2] = {
    {kIn, kHid, 0, kIn * kHid},
    {2 * kHid, kHid, 2112, 2112 + 2 * kHid * kHid},
    {2 * kHid, kHid, 2112 + 8256, 2112 + 8256 + 2 * kHid * kHid},
    {2 * kHid, kHid, 2112 + 2 * 8256, 2112 + 2 * 8256 + 2 * kHid * kHid},
    {2 * kHid, kHid, 2112 + 3 * 8256, 2112 + 3 * 8256 + 2 * kHid * kHid},
    {kHid, kOut, 2112 + 4 * 8256, 2112 + 4 * 8256 + kHid * kOut},
};
static_assert(2112 + 4 * 8256 + kHid * kOut + kOut == BD_NET_PARAMS, "parameter count");

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 as_bf8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }
__device__ __forceinline__ uint4 pack8(const float* f) {
  uint4 o;
  o.x = (uint32_t)f2bf(f[0]) | ((uint32_t)f2bf(f[1]) << 16);
  o.y = (uint32_t)f2bf(f[2]) | ((uint32_t)f2bf(f[3]) << 16);
  o.z = (uint32_t)f2bf(f[4]) | ((uint32_t)f2bf(f[5]) << 16);
  o.w = (uint32_t)f2bf(f[6]) | ((uint32_t)f2bf(f[7]) << 16);
  return o;
}
__device__ __forceinline__ void unpack8(uint4 v, float* f) {
  f[0] = __uint_as_float(v.x << 16);
  f[1] = __uint_as_float(v.x & 0xFFFF0000u);
  f[2] = __uint_as_float(v.y << 16);
  f[3] = __uint_as_float(v.y & 0xFFFF0000u);
  f[4] = __uint_as_float(v.z << 16);
  f[5] = __uint_as_float(v.z & 0xFFFF0000u);
  f[6] = __uint_as_float(v.w << 16);
  f[7] = __uint_as_float(v.w & 0xFFFF0000u);
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// lattice coordinates of level node n
__device__ __forceinline__ void lv_ij(const Geo& g, int n, int& l, int& i, int& j) {
  l = geo_layer(g, n);
  const int r = n - g.off[l];
  j = r / g.nx[l];
  i = r - j * g.nx[l];
}

// the up to 4 lattice neighbours of n in its layer (absent ones: -1)
__device__ __forceinline__ int lattice_nbrs(const Geo& g, int n, int nb[4]) {
  int l, i, j;
  lv_ij(g, n, l, i, j);
  const int nx = g.nx[l], ny = g.ny[l];
  nb[0] = i > 0 ? n - 1 : -1;
  nb[1] = i + 1 < nx ? n + 1 : -1;
  nb[2] = j > 0 ? n - nx : -1;
  nb[3] = j + 1 < ny ? n + nx : -1;
  return (nb[0] >= 0) + (nb[1] >= 0) + (nb[2] >= 0) + (nb[3] >= 0);
}

// ---------------------------------------------------------------- features

struct FeatArgs {
  const float* op;  // level operator, 78 planes
  const float* b;   // the Newton right-hand side restricted to the level, 3 planes
  const uint8_t* fixed;
  int N, Np;
  Geo geo;
  double* partials;
  unsigned int* ticket;
  double* scale;  // [0]: mean trace / 3 of the free rows, [1]: rms of b over them
  uint16_t* x0;   // [Np][kIn]
};

// mean diagonal of the free rows (the operator's scale) and the rms of b, deterministic
__global__ __launch_bounds__(1024) void k_diag_scale(FeatArgs A) {
  double v[3] = {0, 0, 0};
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < A.N; n += gridDim.x * blockDim.x) {
    if (A.fixed[n]) continue;
    v[0] += ((double)A.op[n] + (double)A.op[3 * A.Np + n] + (double)A.op[5 * A.Np + n]) / 3.0;
    v[1] += 1.0;
    for (int q = 0; q < 3; q++) v[2] += (double)A.b[q * A.Np + n] * (double)A.b[q * A.Np + n];
  }
  double tot[3];
  if (grid_sum<3>(v, A.partials, A.ticket, tot)) {
    A.scale[0] = tot[1] > 0 ? tot[0] / tot[1] : 1.0;
    A.scale[1] = (tot[1] > 0 && tot[2] > 0) ? sqrt(tot[2] / (3.0 * tot[1])) : 1.0;
  }
}

// per node: u, v in [0, 1], sin / cos of pi u, pi v, 2 pi u, 2 pi v, fixed flag, asinh of the
// diagonal block over the operator scale (xx xy xz yy yz zz), b over its rms; features 20..31 are zero
__global__ __launch_bounds__(256) void k_features(FeatArgs A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= A.Np) return;
  float f[kIn];
#pragma unroll
  for (int q = 0; q < kIn; q++) f[q] = 0.f;
  if (n < A.N) {
    int l, i, j;
    lv_ij(A.geo, n, l, i, j);
    const float u = A.geo.nx[l] > 1 ? (float)i / (float)(A.geo.nx[l] - 1) : 0.f;
    const float v = A.geo.ny[l] > 1 ? (float)j / (float)(A.geo.ny[l] - 1) : 0.f;
    const float pi = 3.14159265358979f;
    f[0] = u;
    f[1] = v;
    f[2] = sinf(pi * u);
    f[3] = sinf(pi * v);
    f[4] = cosf(pi * u);
    f[5] = cosf(pi * v);
    f[6] = sinf(2.f * pi * u);
    f[7] = sinf(2.f * pi * v);
    f[8] = cosf(2.f * pi * u);
    f[9] = cosf(2.f * pi * v);
    f[10] = A.fixed[n] ? 1.f : 0.f;
    const float inv = (float)(1.0 / A.scale[0]);
    for (int q = 0; q < 6; q++) f[11 + q] = A.fixed[n] ? 0.f : asinhf(A.op[q * A.Np + n] * inv);
    const float binv = (float)(1.0 / A.scale[1]);
    for (int q = 0; q < 3; q++) f[17 + q] = A.fixed[n] ? 0.f : A.b[q * A.Np + n] * binv;
  }
  uint4* o = reinterpret_cast<uint4*>(A.x0 + (size_t)n * kIn);
#pragma unroll
  for (int c = 0; c < kIn / 8; c++) o[c] = pack8(f + 8 * c);
}

// ---------------------------------------------------------------- forward layers

enum { MODE_ENC = 0, MODE_MP = 1, MODE_DEC = 2 };

struct FwdArgs {
  int N, Np;
  Geo geo;
  const uint8_t* fixed;
  const uint16_t* in;   // [Np][K_in] (encoder: x0, K 32; MP: h, K 64; decoder: h, K 64)
  const uint16_t* wt;   // [out][in] bf16
  const float* bias;    // FP32 master bias
  uint16_t* h_out;      // [Np][64]
  uint16_t* apre;       // [Np][64] pre-activations (ReLU masks of the backward)
  uint16_t* mean;       // [Np][64] neighbour means (MP)
  float* z;             // decoder: kNetK * 3 planes of Np
};

template <int MODE>
__global__ __launch_bounds__(256) void k_net_fwd(FwdArgs A) {
  constexpr int K = MODE == MODE_ENC ? kIn : MODE == MODE_MP ? 2 * kHid : kHid;
  constexpr int KIN = MODE == MODE_ENC ? kIn : kHid;  // row length of `in`
  constexpr int NO = MODE == MODE_DEC ? kOut : kHid;
  constexpr int LDW = K + 8;  // padded LDS row (bf16) against bank conflicts
  __shared__ uint16_t wt[NO * LDW];
  for (int t = threadIdx.x; t < NO * K / 8; t += blockDim.x) {
    const int o = t / (K / 8), c = t % (K / 8);
    *reinterpret_cast<uint4*>(&wt[o * LDW + 8 * c]) = reinterpret_cast<const uint4*>(A.wt)[t];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lo = lane & 15, hi = lane >> 4;
  const int n0 = (blockIdx.x * 4 + wave) * 16;
  if (n0 >= A.N) return;
  const int n = n0 + lo;  // < Np: Np is a multiple of 64
  f32x4 acc[NO / 16];
#pragma unroll
  for (int t = 0; t < NO / 16; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < K / 32; ks++) {
    uint4 av;
    if (MODE == MODE_MP && ks >= 2) {
      // mean of the lattice neighbours' features (ks - 2) * 32 + 8 hi .. + 7
      const int fo = (ks - 2) * 32 + 8 * hi;
      float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      int deg = 0;
      if (n < A.N) {
        int nb[4];
        deg = lattice_nbrs(A.geo, n, nb);
        for (int q = 0; q < 4; q++) {
          if (nb[q] < 0) continue;
          float t[8];
          unpack8(*reinterpret_cast<const uint4*>(A.in + (size_t)nb[q] * KIN + fo), t);
#pragma unroll
          for (int e = 0; e < 8; e++) s[e] += t[e];
        }
      }
      const float w = deg > 0 ? 1.f / (float)deg : 0.f;
#pragma unroll
      for (int e = 0; e < 8; e++) s[e] *= w;
      av = pack8(s);
      if (n < A.N) *reinterpret_cast<uint4*>(A.mean + (size_t)n * kHid + fo) = av;
    } else {
      av = *reinterpret_cast<const uint4*>(A.in + (size_t)n * KIN + ks * 32 + 8 * hi);
    }
    const bf16x8 a = as_bf8(av);
#pragma unroll
    for (int t = 0; t < NO / 16; t++) {
      const bf16x8 b = as_bf8(*reinterpret_cast<const uint4*>(&wt[(t * 16 + lo) * LDW + ks * 32 + 8 * hi]));
      acc[t] = mfma(a, b, acc[t]);
    }
  }
  // lane holds rows 4 hi + r (nodes) of column lo (output) of every 16 x 16 tile
#pragma unroll
  for (int t = 0; t < NO / 16; t++) {
    const int o = t * 16 + lo;
    const float bo = A.bias[o];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int node = n0 + 4 * hi + r;
      if (node >= A.N) continue;
      const float a = acc[t][r] + bo;
      if (MODE == MODE_DEC) {
        if (o < 3 * BD_NET_FIELDS) A.z[(size_t)o * A.Np + node] = A.fixed[node] ? 0.f : a;
      } else {
        const float base = MODE == MODE_MP ? bf2f(A.in[(size_t)node * KIN + o]) : 0.f;
        A.apre[(size_t)node * kHid + o] = f2bf(a);
        A.h_out[(size_t)node * kHid + o] = f2bf(base + fmaxf(a, 0.f));
      }
    }
  }
}

// ---------------------------------------------------------------- subspace algebra

struct SubArgs {
  const float* op;  // level operator, 78 planes
  const uint8_t* fixed;
  int N, Np;
  Geo geo;
  const float* z;   // kNetK fields x 3 planes
  float* az;        // A z, same layout
};

__device__ __forceinline__ void ld_blk(const float* op, int Np, int base, int row, float o[6]) {
#pragma unroll
  for (int q = 0; q < 6; q++) o[q] = op[(base + q) * Np + row];
}

// A_l Z for all fields at once: each coupling block is loaded once and applied to every field. The
// products and sums are FP64: for smooth fields a row's terms cancel to a small fraction of their
// size, which FP32 accumulation would not resolve (the Gram matrix and the loss depend on it).
__global__ __launch_bounds__(256) void k_block_apply(SubArgs A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= A.N) return;
  const int Np = A.Np;
  double y[BD_NET_FIELDS][3];
#pragma unroll
  for (int k = 0; k < BD_NET_FIELDS; k++) y[k][0] = y[k][1] = y[k][2] = 0.0;
  if (!A.fixed[n]) {
    int l, i, j;
    lv_ij(A.geo, n, l, i, j);
    const int nx = A.geo.nx[l], ny = A.geo.ny[l];
    auto add = [&](const float b[6], int m) {
#pragma unroll
      for (int k = 0; k < BD_NET_FIELDS; k++) {
        const float* zk = A.z + (size_t)3 * k * Np;
        const double x0 = zk[m], x1 = zk[Np + m], x2 = zk[2 * Np + m];
        y[k][0] += (double)b[0] * x0 + (double)b[1] * x1 + (double)b[2] * x2;
        y[k][1] += (double)b[1] * x0 + (double)b[3] * x1 + (double)b[4] * x2;
        y[k][2] += (double)b[2] * x0 + (double)b[4] * x1 + (double)b[5] * x2;
      }
    };
    float b[6];
    ld_blk(A.op, Np, 0, n, b);
    add(b, n);
    for (int c = 1; c < kCoarseBlocks; c++) {
      const int di = coarse_di(c), dj = coarse_dj(c);
      if (i + di >= 0 && i + di < nx && j + dj < ny) {
        ld_blk(A.op, Np, 6 * c, n, b);
        add(b, n + dj * nx + di);
      }
      if (i - di >= 0 && i - di < nx && j - dj >= 0) {
        const int m = n - dj * nx - di;
        ld_blk(A.op, Np, 6 * c, m, b);
        add(b, m);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < BD_NET_FIELDS; k++)
#pragma unroll
    for (int q = 0; q < 3; q++) A.az[(size_t)(3 * k + q) * Np + n] = (float)y[k][q];
}

struct GramArgs {
  int N, Np;
  const uint8_t* fixed;
  const float* z;
  const float* az;
  const float* b[kTargets];  // target right-hand sides (3 planes each) or null
  double* partials;
  unsigned int* ticket;
  double* G;   // [k][k]
  double* c;   // [kTargets][k]
  int row;
};

// row `row` of the Gram matrix Z^T A Z and of c_t = Z^T b_t (one launch per row)
__global__ __launch_bounds__(1024) void k_gram_row(GramArgs A) {
  constexpr int NV = BD_NET_FIELDS + kTargets;
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) v[q] = 0;
  const int Np = A.Np;
  const float* za = A.z + (size_t)3 * A.row * Np;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < A.N; n += gridDim.x * blockDim.x) {
    if (A.fixed[n]) continue;
    const double x0 = za[n], x1 = za[Np + n], x2 = za[2 * Np + n];
#pragma unroll
    for (int k = 0; k < BD_NET_FIELDS; k++) {
      const float* a = A.az + (size_t)3 * k * Np;
      v[k] += x0 * a[n] + x1 * a[Np + n] + x2 * a[2 * Np + n];
    }
#pragma unroll
    for (int t = 0; t < kTargets; t++)
      if (A.b[t]) v[BD_NET_FIELDS + t] += x0 * A.b[t][n] + x1 * A.b[t][Np + n] + x2 * A.b[t][2 * Np + n];
  }
  double tot[NV];
  if (grid_sum<NV>(v, A.partials, A.ticket, tot)) {
    for (int k = 0; k < BD_NET_FIELDS; k++) A.G[A.row * BD_NET_FIELDS + k] = tot[k];
    for (int t = 0; t < kTargets; t++) A.c[t * BD_NET_FIELDS + A.row] = tot[BD_NET_FIELDS + t];
  }
}

// The pseudo-inverse of the symmetric Gram matrix G in FP64: Jacobi eigen-decomposition, then the
// eigenvalues below kPinvTol times the largest dropped. Fields that are zero or (nearly) dependent
// on the others then add nothing, and P Z G^+ (P Z)^T A stays an A-orthogonal projection.
// One wave: lane l holds entry (l / 8, l % 8) of the matrix and of the rotations' product in
// registers (a single thread kept both 8 x 8 arrays in scratch memory, 320 us a call). Each cyclic
// rotation (p, q) reads its three pivots by lane shuffles and the partner entries of rows and columns
// p and q by lane permutes, in the sequential algorithm's order (columns, then rows, then V).
constexpr double kPinvTol = 1e-7;
__device__ double wave_sum_f64(double x) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  return __shfl(x, 0, 64);
}
__device__ void inverse_spd(const double* G, double* inv) {
  constexpr int K = BD_NET_FIELDS;
  static_assert(K * K == 64, "one matrix entry per lane of a wave");
  const int l = threadIdx.x, i = l / K, j = l % K;
  double a = 0.5 * (G[i * K + j] + G[j * K + i]);
  double v = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 50; sweep++) {
    const double tot = wave_sum_f64(a * a), off = wave_sum_f64(i != j ? a * a : 0.0);
    if (!(off > 1e-30 * tot)) break;
    for (int p = 0; p < K - 1; p++)
      for (int q = p + 1; q < K; q++) {
        const double apq = __shfl(a, p * K + q, 64);
        if (apq == 0.0) continue;
        const double app = __shfl(a, p * K + p, 64), aqq = __shfl(a, q * K + q, 64);
        const double th = (aqq - app) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        const int col = j == p ? i * K + q : (j == q ? i * K + p : l);
        double o = __shfl(a, col, 64);
        if (j == p) a = c * a - sn * o;
        else if (j == q) a = sn * o + c * a;
        o = __shfl(a, i == p ? q * K + j : (i == q ? p * K + j : l), 64);
        if (i == p) a = c * a - sn * o;
        else if (i == q) a = sn * o + c * a;
        o = __shfl(v, col, 64);
        if (j == p) v = c * v - sn * o;
        else if (j == q) v = sn * o + c * v;
      }
  }
  double lmax = i == j ? a : 0.0;
  for (int off = 32; off > 0; off >>= 1) lmax = fmax(lmax, __shfl_down(lmax, off, 64));
  lmax = __shfl(lmax, 0, 64);
  const double d = (i == j && a > kPinvTol * lmax && a > 0) ? 1.0 / a : 0.0;
  double s = 0;
  for (int k = 0; k < K; k++) s += __shfl(v, i * K + k, 64) * __shfl(d, k * K + k, 64) * __shfl(v, j * K + k, 64);
  inv[l] = s;
}

struct LossScal {
  double G[BD_NET_FIELDS * BD_NET_FIELDS];
  double Ginv[BD_NET_FIELDS * BD_NET_FIELDS];
  double c[kTargets * BD_NET_FIELDS];
  double y[kTargets * BD_NET_FIELDS];
  double coef[kTargets];  // -2 w_t / s_t (0: target unused)
  double loss;            // this sample's loss
  double loss_sum;        // summed over the samples of an epoch
  double scale;
  unsigned int ticket;
  int pad;
};

// the inverse (one wave), then the coefficients y_t and the sample's loss (its first lane)
__global__ __launch_bounds__(64) void k_solve_small(LossScal* L, const double* s_t, int ntargets, int mask) {
  inverse_spd(L->G, L->Ginv);
  __syncthreads();
  if (threadIdx.x != 0) return;
  double loss = 0, wsum = 0;
  for (int t = 0; t < kTargets; t++) {
    const bool on = t < ntargets && s_t && s_t[t] > 0 && ((mask >> t) & 1);
    double cy = 0;
    for (int a = 0; a < BD_NET_FIELDS; a++) {
      double y = 0;
      for (int b = 0; b < BD_NET_FIELDS; b++) y += L->Ginv[a * BD_NET_FIELDS + b] * L->c[t * BD_NET_FIELDS + b];
      L->y[t * BD_NET_FIELDS + a] = on ? y : 0.0;
      cy += y * L->c[t * BD_NET_FIELDS + a];
    }
    if (on) {
      wsum += 1.0;
      loss += 1.0 - cy / s_t[t];
    }
  }
  for (int t = 0; t < kTargets; t++) {
    const bool on = t < ntargets && s_t && s_t[t] > 0 && ((mask >> t) & 1);
    L->coef[t] = on ? -2.0 / (wsum * s_t[t]) : 0.0;
  }
  L->loss = wsum > 0 ? loss / wsum : 0.0;
  L->loss_sum += L->loss;
}

// dL/dZ as the decoder's output gradient, node-major bf16 [Np][32] (columns 24..31 zero)
struct DzArgs {
  int N, Np;
  const uint8_t* fixed;
  const float* az;
  const float* b[kTargets];
  const LossScal* L;
  uint16_t* dz;
};

__global__ __launch_bounds__(256) void k_dz(DzArgs A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= A.Np) return;
  float g[kOut];
#pragma unroll
  for (int o = 0; o < kOut; o++) g[o] = 0.f;
  if (n < A.N && !A.fixed[n]) {
    const int Np = A.Np;
    for (int t = 0; t < kTargets; t++) {
      const double cf = A.L->coef[t];
      if (cf == 0.0 || !A.b[t]) continue;
      const double* y = A.L->y + t * BD_NET_FIELDS;
      double r[3];
      for (int q = 0; q < 3; q++) {
        double s = A.b[t][q * Np + n];
        for (int k = 0; k < BD_NET_FIELDS; k++) s -= A.az[(size_t)(3 * k + q) * Np + n] * y[k];
        r[q] = s;
      }
      for (int k = 0; k < BD_NET_FIELDS; k++)
        for (int q = 0; q < 3; q++) g[3 * k + q] += (float)(cf * r[q] * y[k]);
    }
  }
  uint4* o = reinterpret_cast<uint4*>(A.dz + (size_t)n * kOut);
#pragma unroll
  for (int c = 0; c < kOut / 8; c++) o[c] = pack8(g + 8 * c);
}

// ---------------------------------------------------------------- backward layers

// weight-gradient partials of one layer over a block's 64 nodes: P[i][o] = sum_n U[n][i] D[n][o] on
// the matrix cores (U and D staged transposed in LDS so both operands read 16 contiguous bytes), and
// the bias partial sum_n D[n][o]. U = [u0 | u1] (u1 only for the message-passing layers).
struct WgradArgs {
  int N, Np;
  const uint16_t* u0;  // [Np][ld0]
  const uint16_t* u1;  // [Np][64] or null
  int ld0;
  const uint16_t* d;   // [Np][NOUT]
  float* partials;     // [blocks][(KIN + 1) * NOUT]
};

template <int KIN, int NOUT>
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs A) {
  constexpr int LD = kNodesPerBlock + 8;
  __shared__ uint16_t ut[KIN * LD];
  __shared__ uint16_t dt[NOUT * LD];
  const int n0 = blockIdx.x * kNodesPerBlock;
  // stage: thread t takes node t % 64 and a feature chunk
  for (int t = threadIdx.x; t < kNodesPerBlock * (KIN / 8); t += blockDim.x) {
    const int nn = t % kNodesPerBlock, c = t / kNodesPerBlock;
    const int n = n0 + nn;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (n < A.N) {
      const int f = 8 * c;
      v = f < A.ld0 ? *reinterpret_cast<const uint4*>(A.u0 + (size_t)n * A.ld0 + f)
                    : *reinterpret_cast<const uint4*>(A.u1 + (size_t)n * kHid + f - A.ld0);
    }
    const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
    for (int q = 0; q < 8; q++) ut[(8 * c + q) * LD + nn] = e[q];
  }
  for (int t = threadIdx.x; t < kNodesPerBlock * (NOUT / 8); t += blockDim.x) {
    const int nn = t % kNodesPerBlock, c = t / kNodesPerBlock;
    const int n = n0 + nn;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (n < A.N) v = *reinterpret_cast<const uint4*>(A.d + (size_t)n * NOUT + 8 * c);
    const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
    for (int q = 0; q < 8; q++) dt[(8 * c + q) * LD + nn] = e[q];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lo = lane & 15, hi = lane >> 4;
  constexpr int TI = KIN / 16, TO = NOUT / 16, TILES = TI * TO;
  float* out = A.partials + (size_t)blockIdx.x * (KIN + 1) * NOUT;
  for (int tile = wave; tile < TILES; tile += 4) {
    const int ti = tile / TO, to = tile % TO;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < kNodesPerBlock / 32; ks++) {
      const bf16x8 a = as_bf8(*reinterpret_cast<const uint4*>(&ut[(ti * 16 + lo) * LD + ks * 32 + 8 * hi]));
      const bf16x8 b = as_bf8(*reinterpret_cast<const uint4*>(&dt[(to * 16 + lo) * LD + ks * 32 + 8 * hi]));
      acc = mfma(a, b, acc);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) out[(ti * 16 + 4 * hi + r) * NOUT + to * 16 + lo] = acc[r];
  }
  for (int o = threadIdx.x; o < NOUT; o += blockDim.x) {
    float s = 0.f;
    for (int q = 0; q < kNodesPerBlock; q++) s += bf2f(dt[o * LD + q]);
    out[KIN * NOUT + o] = s;
  }
}

// grad[w .. w + KIN NOUT) and grad[b .. b + NOUT) (+)= sum over blocks of the partials, in block order
__global__ __launch_bounds__(256) void k_reduce_partials(const float* partials, int blocks, int entries, int kin, int nout,
                                                         float* grad, int w_off, int b_off) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= entries) return;
  float s = 0.f;
  for (int b = 0; b < blocks; b++) s += partials[(size_t)b * entries + e];
  if (e < kin * nout)
    grad[w_off + e] += s;
  else
    grad[b_off + e - kin * nout] += s;
}

// input gradient dU[n][i] = sum_o D[n][o] W[i][o] (W in-major bf16), FP32 node-major [Np][KIN]
struct IgradArgs {
  int N, Np;
  const uint16_t* d;   // [Np][NOUT]
  const uint16_t* wn;  // [KIN][NOUT]
  float* du;           // [Np][KIN]
};

template <int NOUT, int KIN>
__global__ __launch_bounds__(256) void k_igrad(IgradArgs A) {
  constexpr int LDW = NOUT + 8;
  __shared__ uint16_t wn[KIN * LDW];
  for (int t = threadIdx.x; t < KIN * NOUT / 8; t += blockDim.x) {
    const int i = t / (NOUT / 8), c = t % (NOUT / 8);
    *reinterpret_cast<uint4*>(&wn[i * LDW + 8 * c]) = reinterpret_cast<const uint4*>(A.wn)[t];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lo = lane & 15, hi = lane >> 4;
  const int n0 = (blockIdx.x * 4 + wave) * 16;
  if (n0 >= A.N) return;
  const int n = n0 + lo;
  f32x4 acc[KIN / 16];
#pragma unroll
  for (int t = 0; t < KIN / 16; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < NOUT / 32; ks++) {
    const bf16x8 a = as_bf8(*reinterpret_cast<const uint4*>(A.d + (size_t)n * NOUT + ks * 32 + 8 * hi));
#pragma unroll
    for (int t = 0; t < KIN / 16; t++) {
      const bf16x8 b = as_bf8(*reinterpret_cast<const uint4*>(&wn[(t * 16 + lo) * LDW + ks * 32 + 8 * hi]));
      acc[t] = mfma(a, b, acc[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < KIN / 16; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int node = n0 + 4 * hi + r;
      if (node < A.N) A.du[(size_t)node * KIN + t * 16 + lo] = acc[t][r];
    }
}

// Back through one message-passing layer's input and the ReLU of the layer below:
//   dh_prev[n] = dh[n] + dU[n][0:64] + sum over neighbours m of dU[m][64:128] / deg(m)
//   D[n] = dh_prev[n] * (a_prev[n] > 0)  (bf16, the next weight gradient's operand)
// With du == null it only masks (the decoder's output into the last layer).
struct BackArgs {
  int N, Np;
  Geo geo;
  const float* dh;       // [Np][64]
  const float* du;       // [Np][128] or null
  const uint16_t* apre;  // [Np][64] pre-activations of the layer below
  float* dh_prev;        // [Np][64] (may alias nothing)
  uint16_t* d;           // [Np][64]
};

__global__ __launch_bounds__(256) void k_mp_back(BackArgs A) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = t >> 3, c = t & 7;  // 8 threads per node, 8 features each
  if (n >= A.Np) return;
  float g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (n < A.N) {
    for (int e = 0; e < 8; e++) g[e] = A.dh[(size_t)n * kHid + 8 * c + e];
    if (A.du) {
      for (int e = 0; e < 8; e++) g[e] += A.du[(size_t)n * 2 * kHid + 8 * c + e];
      int nb[4];
      lattice_nbrs(A.geo, n, nb);
      for (int q = 0; q < 4; q++) {
        const int m = nb[q];
        if (m < 0) continue;
        int mb[4];
        const float w = 1.f / (float)lattice_nbrs(A.geo, m, mb);
        for (int e = 0; e < 8; e++) g[e] += w * A.du[(size_t)m * 2 * kHid + kHid + 8 * c + e];
      }
    }
  }
  float dm[8];
  float a[8];
  unpack8(*reinterpret_cast<const uint4*>(A.apre + (size_t)n * kHid + 8 * c), a);
  for (int e = 0; e < 8; e++) dm[e] = (n < A.N && a[e] > 0.f) ? g[e] : 0.f;
  if (A.dh_prev)
    for (int e = 0; e < 8; e++) A.dh_prev[(size_t)n * kHid + 8 * c + e] = g[e];
  *reinterpret_cast<uint4*>(A.d + (size_t)n * kHid + 8 * c) = pack8(dm);
}

// Adam on the FP32 master weights, then their bf16 copies in both layouts
__global__ __launch_bounds__(256) void k_adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1,
                                              float b2, float eps, float c1, float c2, float gscale) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gr = g[e] * gscale;
  const float mm = b1 * m[e] + (1.f - b1) * gr;
  const float vv = b2 * v[e] + (1.f - b2) * gr * gr;
  m[e] = mm;
  v[e] = vv;
  w[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

struct PackArgs {
  const float* w;
  uint16_t* wt[kMp + 2];
  uint16_t* wn[kMp + 2];
};

__global__ __launch_bounds__(256) void k_pack_weights(PackArgs A) {
  const int layer = blockIdx.y;
  const LayerDesc L = kLayers[layer];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L.in * L.out) return;
  const int i = e / L.out, o = e % L.out;
  const uint16_t b = f2bf(A.w[L.w + e]);
  A.wn[layer][i * L.out + o] = b;
  A.wt[layer][o * L.in + i] = b;
}

// ---------------------------------------------------------------- the correction in the solve

struct CorrArgs {
  int N, Np;              // net level
  const uint8_t* fixed;
  const float* z;
  float* t;               // the level vector: restricted residual in, correction out
  double* partials;
  unsigned int* ticket;
  LossScal* L;
  const int* done;        // PCG's done flag
  const int* off;         // the learned correction's off switch
};

// c = Z^T t, y = Ginv c (into L->y[0..k))
__global__ __launch_bounds__(1024) void k_corr_dots(CorrArgs A) {
  if (*A.done || *A.off) return;
  double v[BD_NET_FIELDS];
#pragma unroll
  for (int k = 0; k < BD_NET_FIELDS; k++) v[k] = 0;
  const int Np = A.Np;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < A.N; n += gridDim.x * blockDim.x) {
    if (A.fixed[n]) continue;
    const double t0 = A.t[n], t1 = A.t[Np + n], t2 = A.t[2 * Np + n];
#pragma unroll
    for (int k = 0; k < BD_NET_FIELDS; k++) {
      const float* zk = A.z + (size_t)3 * k * Np;
      v[k] += t0 * zk[n] + t1 * zk[Np + n] + t2 * zk[2 * Np + n];
    }
  }
  double tot[BD_NET_FIELDS];
  if (grid_sum<BD_NET_FIELDS>(v, A.partials, A.ticket, tot)) {
    for (int a = 0; a < BD_NET_FIELDS; a++) {
      double y = 0;
      for (int b = 0; b < BD_NET_FIELDS; b++) y += A.L->Ginv[a * BD_NET_FIELDS + b] * tot[b];
      A.L->y[a] = y;
    }
  }
}

// W^T z over the fine level (the deflation's dots), y = -Ginv W^T z into L->y
struct DeflArgs {
  int N, Np;
  const uint8_t* fixed;
  const float* w;  // kNetK fields x 3 planes of the fine level
  const float* z;
  double* partials;
  unsigned int* ticket;
  LossScal* L;
  const int* done;
  const int* off;
};

__global__ __launch_bounds__(1024) void k_defl_dots(DeflArgs A) {
  if (*A.done || *A.off) return;
  double v[BD_NET_FIELDS];
#pragma unroll
  for (int k = 0; k < BD_NET_FIELDS; k++) v[k] = 0;
  const int Np = A.Np;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < A.N; n += gridDim.x * blockDim.x) {
    if (A.fixed[n]) continue;
    const double z0 = A.z[n], z1 = A.z[Np + n], z2 = A.z[2 * Np + n];
#pragma unroll
    for (int k = 0; k < BD_NET_FIELDS; k++) {
      const float* wk = A.w + (size_t)3 * k * Np;
      v[k] += z0 * wk[n] + z1 * wk[Np + n] + z2 * wk[2 * Np + n];
    }
  }
  double tot[BD_NET_FIELDS];
  if (grid_sum<BD_NET_FIELDS>(v, A.partials, A.ticket, tot)) {
    for (int a = 0; a < BD_NET_FIELDS; a++) {
      double y = 0;
      for (int b = 0; b < BD_NET_FIELDS; b++) y += A.L->Ginv[a * BD_NET_FIELDS + b] * tot[b];
      A.L->y[a] = -y;
    }
  }
}

// column `col` of G = (P Z)^T (A P Z) from the restricted column R A P Z_col at the net level
__global__ __launch_bounds__(1024) void k_gram_col(CorrArgs A, int col) {
  double v[BD_NET_FIELDS];
#pragma unroll
  for (int k = 0; k < BD_NET_FIELDS; k++) v[k] = 0;
  const int Np = A.Np;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < A.N; n += gridDim.x * blockDim.x) {
    if (A.fixed[n]) continue;
    const double t0 = A.t[n], t1 = A.t[Np + n], t2 = A.t[2 * Np + n];
#pragma unroll
    for (int k = 0; k < BD_NET_FIELDS; k++) {
      const float* zk = A.z + (size_t)3 * k * Np;
      v[k] += t0 * zk[n] + t1 * zk[Np + n] + t2 * zk[2 * Np + n];
    }
  }
  double tot[BD_NET_FIELDS];
  if (grid_sum<BD_NET_FIELDS>(v, A.partials, A.ticket, tot))
    for (int a = 0; a < BD_NET_FIELDS; a++) A.L->G[a * BD_NET_FIELDS + col] = tot[a];
}

// the symmetric part of G (FP32 rounding leaves it slightly asymmetric) and its inverse (one wave)
__global__ __launch_bounds__(64) void k_gram_inverse(LossScal* L) {
  if (threadIdx.x == 0)
    for (int a = 0; a < BD_NET_FIELDS; a++)
      for (int b = 0; b < a; b++) {
        const double m = 0.5 * (L->G[a * BD_NET_FIELDS + b] + L->G[b * BD_NET_FIELDS + a]);
        L->G[a * BD_NET_FIELDS + b] = L->G[b * BD_NET_FIELDS + a] = m;
      }
  __syncthreads();
  inverse_spd(L->G, L->Ginv);
}

// one field of Z into the level vector t
__global__ __launch_bounds__(256) void k_copy_field(const float* z, int Np, int k, float* t) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= Np) return;
  for (int q = 0; q < 3; q++) t[q * Np + n] = z[(size_t)(3 * k + q) * Np + n];
}

// t = Z y
__global__ __launch_bounds__(256) void k_corr_combine(CorrArgs A) {
  if (*A.done || *A.off) return;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= A.N) return;
  const int Np = A.Np;
  float s[3] = {0.f, 0.f, 0.f};
  if (!A.fixed[n])
    for (int k = 0; k < BD_NET_FIELDS; k++) {
      const float y = (float)A.L->y[k];
      const float* zk = A.z + (size_t)3 * k * Np;
      s[0] += y * zk[n];
      s[1] += y * zk[Np + n];
      s[2] += y * zk[2 * Np + n];
    }
  for (int q = 0; q < 3; q++) A.t[q * Np + n] = s[q];
}

}  // namespace

// ---------------------------------------------------------------- host side

struct Learn {
  int level = -1;
  int Nl = 0, Nlp = 0, blocks = 0;
  // network
  float* w = nullptr;  // FP32 master weights
  float* grad = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  int adam_t = 0;
  uint16_t* wt[kMp + 2] = {};
  uint16_t* wn[kMp + 2] = {};
  uint16_t* x0 = nullptr;
  uint16_t* h[kMp + 1] = {};
  uint16_t* apre[kMp + 1] = {};
  uint16_t* mean[kMp] = {};
  float* z = nullptr;
  float* az = nullptr;
  float* wf = nullptr;    // W = A P Z on the fine level (kNetK fields x 3 planes)
  float* ftmp = nullptr;  // a fine vector
  double* dummy = nullptr;
  uint16_t* dz = nullptr;
  float* dh[2] = {};
  float* du = nullptr;
  uint16_t* d = nullptr;
  float* partials = nullptr;
  double* rpart = nullptr;  // FP64 reduction partials
  LossScal* L = nullptr;
  double* scale = nullptr;  // [2]
  bool have_z = false;
  // restriction / prolongation chain vectors, levels 1 .. level
  float* chain[kMaxLevels] = {};
  // recording
  int capacity = 0, count = 0;
  std::vector<float*> s_op;
  std::vector<uint8_t*> s_fixed;
  std::vector<float*> s_b;   // kTargets x 3 planes
  double* s_s = nullptr;     // [capacity][kTargets]
  float* rec_d = nullptr;    // fine: the iterate and residual at kSlowIter
  float* rec_r = nullptr;
  bool rec_mid = false;
};

namespace {

template <class T>
T* lalloc(size_t n) {
  void* p = nullptr;
  BD_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  BD_HIP(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
  BD_HIP(hipDeviceSynchronize());
  return static_cast<T*>(p);
}

int pick_level(const Solver& s) {
  if (s.P.net_level > 0) return std::min(s.P.net_level, s.nlev - 1);
  for (int k = 1; k < s.nlev; k++)
    if (s.lv[k].N <= 5000) return k;
  return s.nlev - 1;
}

void init_weights(Learn& G, Solver& s) {
  std::vector<float> w(BD_NET_PARAMS, 0.f);
  std::mt19937 rng(20261204u);
  for (int l = 0; l < kMp + 2; l++) {
    const LayerDesc& L = kLayers[l];
    // the message-passing layers start small (near-identity residual blocks)
    const float a = std::sqrt(6.f / (float)(L.in + L.out)) * (l >= 1 && l <= kMp ? 0.3f : 1.f);
    std::uniform_real_distribution<float> U(-a, a);
    for (int e = 0; e < L.in * L.out; e++) w[L.w + e] = U(rng);
    if (l == 0)
      for (int i = 20; i < kIn; i++)
        for (int o = 0; o < L.out; o++) w[L.w + i * L.out + o] = 0.f;
  }
  BD_HIP(hipMemcpyAsync(G.w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, s.stream));
  BD_HIP(hipStreamSynchronize(s.stream));
}

void pack_weights(Learn& G, Solver& s) {
  PackArgs A{G.w, {}, {}};
  for (int l = 0; l < kMp + 2; l++) {
    A.wt[l] = G.wt[l];
    A.wn[l] = G.wn[l];
  }
  k_pack_weights<<<dim3(grid_for(2 * kHid * kHid, 256), kMp + 2), 256, 0, s.stream>>>(A);
}

Geo level_geo(const Solver& s, int k) { return s.lv[k].geo; }

// the network on the features of an operator at the net level and a restricted rhs: Z (kNetK fields)
void forward(Learn& G, Solver& s, const float* op, const uint8_t* fixed, const float* b) {
  const Geo g = level_geo(s, G.level);
  FeatArgs F{op, b, fixed, G.Nl, G.Nlp, g, G.rpart, &G.L->ticket, G.scale, G.x0};
  k_diag_scale<<<std::min(kDotBlocks, grid_for(G.Nl, 1024)), 1024, 0, s.stream>>>(F);
  k_features<<<grid_for(G.Nlp, 256), 256, 0, s.stream>>>(F);
  FwdArgs A{G.Nl, G.Nlp, g, fixed, G.x0, G.wt[0], G.w + kLayers[0].b, G.h[0], G.apre[0], nullptr, nullptr};
  k_net_fwd<MODE_ENC><<<G.blocks, 256, 0, s.stream>>>(A);
  for (int l = 1; l <= kMp; l++) {
    A = FwdArgs{G.Nl, G.Nlp, g, fixed, G.h[l - 1], G.wt[l], G.w + kLayers[l].b, G.h[l], G.apre[l], G.mean[l - 1], nullptr};
    k_net_fwd<MODE_MP><<<G.blocks, 256, 0, s.stream>>>(A);
  }
  A = FwdArgs{G.Nl, G.Nlp, g, fixed, G.h[kMp], G.wt[kMp + 1], G.w + kLayers[kMp + 1].b, nullptr, nullptr, nullptr, G.z};
  k_net_fwd<MODE_DEC><<<G.blocks, 256, 0, s.stream>>>(A);
}

// A Z, the Gram matrix and c_t for targets b, its inverse, y_t and the loss
void subspace(Learn& G, Solver& s, const float* op, const uint8_t* fixed, const float* const* b, const double* s_t,
              int mask = 3) {
  const Geo g = level_geo(s, G.level);
  SubArgs S{op, fixed, G.Nl, G.Nlp, g, G.z, G.az};
  k_block_apply<<<grid_for(G.Nl, 256), 256, 0, s.stream>>>(S);
  for (int row = 0; row < BD_NET_FIELDS; row++) {
    GramArgs A{G.Nl, G.Nlp, fixed, G.z, G.az, {b ? b[0] : nullptr, b ? b[1] : nullptr}, G.rpart, &G.L->ticket, G.L->G,
               G.L->c, row};
    k_gram_row<<<std::min(kDotBlocks, grid_for(G.Nl, 1024)), 1024, 0, s.stream>>>(A);
  }
  k_solve_small<<<1, 64, 0, s.stream>>>(G.L, s_t, b ? kTargets : 0, mask);
}

void backward(Learn& G, Solver& s, const uint8_t* fixed, const float* const* b) {
  const Geo g = level_geo(s, G.level);
  DzArgs D{G.Nl, G.Nlp, fixed, G.az, {b[0], b[1]}, G.L, G.dz};
  k_dz<<<grid_for(G.Nlp, 256), 256, 0, s.stream>>>(D);
  const int nb = grid_for(G.Nl, kNodesPerBlock);
  auto reduce = [&](int layer) {
    const LayerDesc& L = kLayers[layer];
    const int entries = (L.in + 1) * L.out;
    k_reduce_partials<<<grid_for(entries, 256), 256, 0, s.stream>>>(G.partials, nb, entries, L.in, L.out, G.grad, L.w, L.b);
  };
  // decoder
  {
    WgradArgs W{G.Nl, G.Nlp, G.h[kMp], nullptr, kHid, G.dz, G.partials};
    k_wgrad<kHid, kOut><<<nb, 256, 0, s.stream>>>(W);
    reduce(kMp + 1);
    IgradArgs I{G.Nl, G.Nlp, G.dz, G.wn[kMp + 1], G.dh[0]};
    k_igrad<kOut, kHid><<<G.blocks, 256, 0, s.stream>>>(I);
    BackArgs B{G.Nl, G.Nlp, g, G.dh[0], nullptr, G.apre[kMp], nullptr, G.d};
    k_mp_back<<<grid_for(8 * G.Nlp, 256), 256, 0, s.stream>>>(B);
  }
  int cur = 0;
  for (int l = kMp; l >= 1; l--) {
    WgradArgs W{G.Nl, G.Nlp, G.h[l - 1], G.mean[l - 1], kHid, G.d, G.partials};
    k_wgrad<2 * kHid, kHid><<<nb, 256, 0, s.stream>>>(W);
    reduce(l);
    IgradArgs I{G.Nl, G.Nlp, G.d, G.wn[l], G.du};
    k_igrad<kHid, 2 * kHid><<<G.blocks, 256, 0, s.stream>>>(I);
    BackArgs B{G.Nl, G.Nlp, g, G.dh[cur], G.du, G.apre[l - 1], G.dh[1 - cur], G.d};
    k_mp_back<<<grid_for(8 * G.Nlp, 256), 256, 0, s.stream>>>(B);
    cur = 1 - cur;
  }
  WgradArgs W{G.Nl, G.Nlp, G.x0, nullptr, kIn, G.d, G.partials};
  k_wgrad<kIn, kHid><<<nb, 256, 0, s.stream>>>(W);
  reduce(0);
}

// t at the net level from a fine vector r: restriction chain (level k: chain[k])
void restrict_to(Learn& G, Solver& s, const float* r, const int* gate) {
  const float* src = r;
  for (int k = 0; k < G.level; k++) {
    launch_restrict_vec(s, k, src, G.chain[k + 1], gate);
    src = G.chain[k + 1];
  }
}

// dst (+)= P chain of the net-level vector chain[level]
void prolong_from(Learn& G, Solver& s, float* dst, const int* gate) {
  for (int k = G.level - 1; k >= 0; k--) launch_prolong_vec(s, k, G.chain[k + 1], k == 0 ? dst : G.chain[k], k > 0, gate);
}

}  // namespace

void learn_free(Solver& s) {
  Learn* G = s.learn;
  if (!G) return;
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(G->w);
  f(G->grad);
  f(G->adam_m);
  f(G->adam_v);
  for (int l = 0; l < kMp + 2; l++) {
    f(G->wt[l]);
    f(G->wn[l]);
  }
  f(G->x0);
  for (int l = 0; l <= kMp; l++) {
    f(G->h[l]);
    f(G->apre[l]);
  }
  for (int l = 0; l < kMp; l++) f(G->mean[l]);
  f(G->z);
  f(G->az);
  f(G->wf);
  f(G->ftmp);
  f(G->dummy);
  f(G->dz);
  f(G->dh[0]);
  f(G->dh[1]);
  f(G->du);
  f(G->d);
  f(G->partials);
  f(G->rpart);
  f(G->L);
  f(G->scale);
  for (int k = 0; k < kMaxLevels; k++) f(G->chain[k]);
  for (auto p : G->s_op) f(p);
  for (auto p : G->s_fixed) f(p);
  for (auto p : G->s_b) f(p);
  f(G->s_s);
  f(G->rec_d);
  f(G->rec_r);
  delete G;
  s.learn = nullptr;
}

void learn_alloc(Solver& s) {
  if (s.learn) return;
  if (s.nlev < 2) throw std::invalid_argument("the learned correction needs a hierarchy of at least two levels");
  Learn* G = new Learn();
  s.learn = G;
  G->level = pick_level(s);
  const Level& L = s.lv[G->level];
  G->Nl = L.N;
  G->Nlp = L.Np;
  G->blocks = grid_for(G->Nl, kNodesPerBlock);
  G->w = lalloc<float>(BD_NET_PARAMS);
  G->grad = lalloc<float>(BD_NET_PARAMS);
  G->adam_m = lalloc<float>(BD_NET_PARAMS);
  G->adam_v = lalloc<float>(BD_NET_PARAMS);
  for (int l = 0; l < kMp + 2; l++) {
    G->wt[l] = lalloc<uint16_t>((size_t)kLayers[l].in * kLayers[l].out);
    G->wn[l] = lalloc<uint16_t>((size_t)kLayers[l].in * kLayers[l].out);
  }
  const size_t Np = G->Nlp;
  G->x0 = lalloc<uint16_t>(Np * kIn);
  for (int l = 0; l <= kMp; l++) {
    G->h[l] = lalloc<uint16_t>(Np * kHid);
    G->apre[l] = lalloc<uint16_t>(Np * kHid);
  }
  for (int l = 0; l < kMp; l++) G->mean[l] = lalloc<uint16_t>(Np * kHid);
  G->z = lalloc<float>(Np * 3 * BD_NET_FIELDS);
  G->wf = lalloc<float>((size_t)s.Np * 3 * BD_NET_FIELDS);
  G->ftmp = lalloc<float>((size_t)s.Np * 3);
  G->dummy = lalloc<double>(1);
  G->az = lalloc<float>(Np * 3 * BD_NET_FIELDS);
  G->dz = lalloc<uint16_t>(Np * kOut);
  G->dh[0] = lalloc<float>(Np * kHid);
  G->dh[1] = lalloc<float>(Np * kHid);
  G->du = lalloc<float>(Np * 2 * kHid);
  G->d = lalloc<uint16_t>(Np * kHid);
  G->partials = lalloc<float>((size_t)grid_for(G->Nl, kNodesPerBlock) * (2 * kHid + 1) * kHid);
  G->rpart = lalloc<double>((size_t)kDotBlocks * (BD_NET_FIELDS + kTargets));
  G->L = lalloc<LossScal>(1);
  G->scale = lalloc<double>(2);
  for (int k = 1; k <= G->level; k++) G->chain[k] = lalloc<float>(3 * (size_t)s.lv[k].Np);
  init_weights(*G, s);
  pack_weights(*G, s);
  BD_HIP(hipStreamSynchronize(s.stream));
}

int learn_level(const Solver& s) { return s.learn ? s.learn->level : -1; }
int learn_padded(const Solver& s) { return s.learn ? s.learn->Nlp : 0; }
int learn_nodes(const Solver& s) { return s.learn ? s.learn->Nl : 0; }

__global__ void k_learn_gate(Scal* sc) { sc->lgate = (sc->done || sc->learn_off) ? 1 : 0; }

// after every hierarchy build while the learned correction is on: Z from the net level's operator
// and the restricted rhs; then per field W = A P Z on the fine level and G = (P Z)^T W (exact for the
// current fine operator), and G's inverse
void learn_predict(Solver& s) {
  Learn& G = *s.learn;
  const Level& L = s.lv[G.level];
  ProfScope ps(s, PC_LEARNED, 0.0);
  restrict_to(G, s, s.rhs, nullptr);
  forward(G, s, L.op, L.fixed, G.chain[G.level]);
  if (s.learned != BD_LEARNED_DEFLATE) {
    // the warm start alone: G = Z^T A_l Z on the net level (the Galerkin operator, FP64 sums)
    subspace(G, s, L.op, L.fixed, nullptr, nullptr);
    G.have_z = true;
    return;
  }
  CorrArgs C{G.Nl, G.Nlp, L.fixed, G.z, G.chain[G.level], G.rpart, &G.L->ticket, G.L, nullptr, nullptr};
  for (int k = 0; k < BD_NET_FIELDS; k++) {
    k_copy_field<<<grid_for(G.Nlp, 256), 256, 0, s.stream>>>(G.z, G.Nlp, k, G.chain[G.level]);
    launch_zero(s, G.ftmp, 3 * (size_t)s.Np);
    prolong_from(G, s, G.ftmp, nullptr);
    float* wk = G.wf + (size_t)3 * k * s.Np;
    launch_apply_f64(s, 0, G.ftmp, wk);
    restrict_to(G, s, wk, nullptr);
    k_gram_col<<<std::min(kDotBlocks, grid_for(G.Nl, 1024)), 1024, 0, s.stream>>>(C, k);
  }
  k_gram_inverse<<<1, 64, 0, s.stream>>>(G.L);
  G.have_z = true;
}

// the deflation of a preconditioned residual: z -= P Z Ginv W^T z (gated like the correction)
void learn_deflate(Solver& s, float* z) {
  Learn& G = *s.learn;
  const int* gate = learn_gate(s);
  ProfScope ps(s, PC_LEARNED, 0.0);
  k_learn_gate<<<1, 1, 0, s.stream>>>(s.scal);
  DeflArgs D{s.N, s.Np, s.pinned, G.wf, z, G.rpart, &G.L->ticket, G.L, pcg_done(s), learn_off(s)};
  k_defl_dots<<<std::min(kDotBlocks, grid_for(s.N, 1024)), 1024, 0, s.stream>>>(D);
  CorrArgs C{G.Nl, G.Nlp, s.lv[G.level].fixed, G.z, G.chain[G.level], G.rpart, &G.L->ticket, G.L, pcg_done(s), learn_off(s)};
  k_corr_combine<<<grid_for(G.Nl, 256), 256, 0, s.stream>>>(C);
  prolong_from(G, s, z, gate);
}

// z += P Z Ginv Z^T R r (gated by PCG's done flag and the correction's off switch)
void learn_correct(Solver& s, const float* r, float* z) {
  Learn& G = *s.learn;
  const int* gate = learn_gate(s);
  ProfScope ps(s, PC_LEARNED, 0.0);
  k_learn_gate<<<1, 1, 0, s.stream>>>(s.scal);
  restrict_to(G, s, r, gate);
  CorrArgs C{G.Nl, G.Nlp, s.lv[G.level].fixed, G.z, G.chain[G.level], G.rpart, &G.L->ticket, G.L, pcg_done(s), learn_off(s)};
  k_corr_dots<<<std::min(kDotBlocks, grid_for(G.Nl, 1024)), 1024, 0, s.stream>>>(C);
  k_corr_combine<<<grid_for(G.Nl, 256), 256, 0, s.stream>>>(C);
  prolong_from(G, s, z, gate);
}

bool learn_ready(const Solver& s) { return s.learn && s.learn->have_z; }

// ---- recording training samples from plain solves

void learn_record_start(Solver& s, int capacity) {
  learn_alloc(s);
  Learn& G = *s.learn;
  for (auto p : G.s_op) (void)hipFree(p);
  for (auto p : G.s_fixed) (void)hipFree(p);
  for (auto p : G.s_b) (void)hipFree(p);
  G.s_op.clear();
  G.s_fixed.clear();
  G.s_b.clear();
  if (G.s_s) (void)hipFree(G.s_s);
  G.s_s = nullptr;
  G.capacity = std::max(0, capacity);
  G.count = 0;
  if (G.capacity == 0) return;
  G.s_s = lalloc<double>((size_t)G.capacity * kTargets);
  if (!G.rec_d) {
    G.rec_d = lalloc<float>(3 * (size_t)s.Np);
    G.rec_r = lalloc<float>(3 * (size_t)s.Np);
  }
}

bool learn_recording(const Solver& s) { return s.learn && s.learn->count < s.learn->capacity; }

__global__ void k_copy_f(const float* a, float* b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_copy_delta(const double* d, float* o, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = (float)d[i];
}

// at PCG start (r = b): the sample's operator, fixed flags and the warm-start target R b
void learn_record_begin(Solver& s, const float* r) {
  Learn& G = *s.learn;
  const Level& L = s.lv[G.level];
  const int c = G.count;
  if ((int)G.s_op.size() <= c) {
    G.s_op.push_back(lalloc<float>((size_t)kCoarseComps * G.Nlp));
    G.s_fixed.push_back(lalloc<uint8_t>(G.Nlp));
    G.s_b.push_back(lalloc<float>((size_t)kTargets * 3 * G.Nlp));
  }
  BD_HIP(hipMemcpyAsync(G.s_op[c], L.op, (size_t)kCoarseComps * G.Nlp * sizeof(float), hipMemcpyDeviceToDevice, s.stream));
  BD_HIP(hipMemcpyAsync(G.s_fixed[c], L.fixed, G.Nlp, hipMemcpyDeviceToDevice, s.stream));
  restrict_to(G, s, r, nullptr);
  k_copy_f<<<grid_for(3 * G.Nlp, 256), 256, 0, s.stream>>>(G.chain[G.level], G.s_b[c], 3 * (size_t)G.Nlp);
  G.rec_mid = false;
}

// at iteration kSlowIter: the iterate and the residual (the slow-mode target is R r_m, with energy
// (d - d_m) . r_m at the end)
void learn_record_mid(Solver& s, const float* r, const double* delta) {
  Learn& G = *s.learn;
  const int c = G.count;
  k_copy_delta<<<grid_for(3 * s.Np, 256), 256, 0, s.stream>>>(delta, G.rec_d, 3 * (size_t)s.Np);
  k_copy_f<<<grid_for(3 * s.Np, 256), 256, 0, s.stream>>>(r, G.rec_r, 3 * (size_t)s.Np);
  restrict_to(G, s, r, nullptr);
  k_copy_f<<<grid_for(3 * G.Nlp, 256), 256, 0, s.stream>>>(G.chain[G.level], G.s_b[c] + 3 * (size_t)G.Nlp, 3 * (size_t)G.Nlp);
  G.rec_mid = true;
}

struct EndArgs {
  int N, Np;
  const uint8_t* fixed;
  const double* delta;
  const float* rhs;
  const float* dm;
  const float* rm;
  int mid;
  double* partials;
  unsigned int* ticket;
  double* out;  // [kTargets]
};

__global__ __launch_bounds__(1024) void k_record_end(EndArgs A) {
  double v[2] = {0, 0};
  const int Np = A.Np;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < A.N; n += gridDim.x * blockDim.x) {
    if (A.fixed[n]) continue;
    for (int q = 0; q < 3; q++) {
      const double d = A.delta[q * Np + n];
      v[0] += d * (double)A.rhs[q * Np + n];
      if (A.mid) v[1] += (d - (double)A.dm[q * Np + n]) * (double)A.rm[q * Np + n];
    }
  }
  double tot[2];
  if (grid_sum<2>(v, A.partials, A.ticket, tot)) {
    A.out[0] = tot[0] > 0 ? tot[0] : 0.0;
    A.out[1] = (A.mid && tot[1] > 0) ? tot[1] : 0.0;
  }
}

// at the end of the solve: the targets' energies; the sample is complete
void learn_record_end(Solver& s, const double* delta, const float* rhs) {
  Learn& G = *s.learn;
  EndArgs E{s.N, s.Np, s.pinned, delta, rhs, G.rec_d, G.rec_r, G.rec_mid ? 1 : 0, G.rpart, &G.L->ticket,
            G.s_s + (size_t)G.count * kTargets};
  k_record_end<<<std::min(kDotBlocks, grid_for(s.N, 1024)), 1024, 0, s.stream>>>(E);
  G.count += 1;
}

int learn_sample_count(const Solver& s) { return s.learn ? s.learn->count : 0; }

// Adam over the recorded samples; losses[e] = mean loss of epoch e (before its updates)
void learn_train(Solver& s, int epochs, double lr, int batch, int targets, double* losses) {
  Learn& G = *s.learn;
  if (G.count == 0) throw std::invalid_argument("no recorded training samples");
  batch = std::max(1, batch);
  std::mt19937 rng(7u + (unsigned)G.adam_t);
  std::vector<int> order(G.count);
  for (int e = 0; e < epochs; e++) {
    for (int q = 0; q < G.count; q++) order[q] = q;
    std::shuffle(order.begin(), order.end(), rng);
    BD_HIP(hipMemsetAsync(&G.L->loss_sum, 0, sizeof(double), s.stream));
    for (int q0 = 0; q0 < G.count; q0 += batch) {
      BD_HIP(hipMemsetAsync(G.grad, 0, BD_NET_PARAMS * sizeof(float), s.stream));
      const int q1 = std::min(G.count, q0 + batch);
      for (int q = q0; q < q1; q++) {
        const int c = order[q];
        const float* b[kTargets] = {G.s_b[c], G.s_b[c] + 3 * (size_t)G.Nlp};
        forward(G, s, G.s_op[c], G.s_fixed[c], b[0]);
        subspace(G, s, G.s_op[c], G.s_fixed[c], b, G.s_s + (size_t)c * kTargets, targets);
        backward(G, s, G.s_fixed[c], b);
      }
      G.adam_t += 1;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - std::pow(b1, (float)G.adam_t), c2 = 1.f - std::pow(b2, (float)G.adam_t);
      k_adam<<<grid_for(BD_NET_PARAMS, 256), 256, 0, s.stream>>>(G.w, G.grad, G.adam_m, G.adam_v, BD_NET_PARAMS, (float)lr,
                                                                    b1, b2, 1e-8f, c1, c2, 1.f / (float)(q1 - q0));
      pack_weights(G, s);
    }
    double sum = 0;
    BD_HIP(hipMemcpyAsync(&sum, &G.L->loss_sum, sizeof(double), hipMemcpyDeviceToHost, s.stream));
    BD_HIP(hipStreamSynchronize(s.stream));
    if (losses) losses[e] = sum / G.count;
  }
  G.have_z = false;  // the next hierarchy build predicts with the new weights
}

// loss and gradient of one sample at the current weights (the gradient check's entry)
void learn_loss_grad(Solver& s, int sample, double* loss, float* grad) {
  Learn& G = *s.learn;
  if (sample < 0 || sample >= G.count) throw std::invalid_argument("no such sample");
  const float* b[kTargets] = {G.s_b[sample], G.s_b[sample] + 3 * (size_t)G.Nlp};
  BD_HIP(hipMemsetAsync(G.grad, 0, BD_NET_PARAMS * sizeof(float), s.stream));
  forward(G, s, G.s_op[sample], G.s_fixed[sample], b[0]);
  subspace(G, s, G.s_op[sample], G.s_fixed[sample], b, G.s_s + (size_t)sample * kTargets);
  backward(G, s, G.s_fixed[sample], b);
  BD_HIP(hipMemcpyAsync(loss, &G.L->loss, sizeof(double), hipMemcpyDeviceToHost, s.stream));
  if (grad) BD_HIP(hipMemcpyAsync(grad, G.grad, BD_NET_PARAMS * sizeof(float), hipMemcpyDeviceToHost, s.stream));
  BD_HIP(hipStreamSynchronize(s.stream));
  G.have_z = false;
}

void learn_get_weights(Solver& s, float* out) {
  BD_HIP(hipMemcpyAsync(out, s.learn->w, BD_NET_PARAMS * sizeof(float), hipMemcpyDeviceToHost, s.stream));
  BD_HIP(hipStreamSynchronize(s.stream));
}

void learn_set_weights(Solver& s, const float* w) {
  BD_HIP(hipMemcpyAsync(s.learn->w, w, BD_NET_PARAMS * sizeof(float), hipMemcpyHostToDevice, s.stream));
  pack_weights(*s.learn, s);
  BD_HIP(hipStreamSynchronize(s.stream));
  s.learn->have_z = false;
}

// a recorded sample's targets at the net level (3 planes of Nlp per target) and energies
void learn_read_sample(Solver& s, int sample, float* b, double* st, float* op, uint8_t* fixed) {
  Learn& G = *s.learn;
  if (sample < 0 || sample >= G.count) throw std::invalid_argument("no such sample");
  if (b) BD_HIP(hipMemcpyAsync(b, G.s_b[sample], (size_t)kTargets * 3 * G.Nlp * sizeof(float), hipMemcpyDeviceToHost, s.stream));
  if (st) BD_HIP(hipMemcpyAsync(st, G.s_s + (size_t)sample * kTargets, kTargets * sizeof(double), hipMemcpyDeviceToHost, s.stream));
  if (op) BD_HIP(hipMemcpyAsync(op, G.s_op[sample], (size_t)kCoarseComps * G.Nlp * sizeof(float), hipMemcpyDeviceToHost, s.stream));
  if (fixed) BD_HIP(hipMemcpyAsync(fixed, G.s_fixed[sample], G.Nlp, hipMemcpyDeviceToHost, s.stream));
  BD_HIP(hipStreamSynchronize(s.stream));
}

// the last prediction's fields (kNetK x 3 planes of Nlp)
void learn_read_z(Solver& s, float* z) {
  BD_HIP(hipMemcpyAsync(z, s.learn->z, (size_t)3 * BD_NET_FIELDS * s.learn->Nlp * sizeof(float), hipMemcpyDeviceToHost, s.stream));
  BD_HIP(hipStreamSynchronize(s.stream));
}

}  // namespace bd

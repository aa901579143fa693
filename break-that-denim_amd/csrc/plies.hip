// Cross-layer contact of a scene of several plies: each crossing node against the crossing nodes of
// the other layers. Restates src/simulation/gpu/wgsl/contact.wgsl:1-164 (hashCount, hashFill,
// detect, terms) with the law of gpu/contact.ts:8-18 (CONTACT) and the pair distances of
// gpu/contact.ts:52-59: half the summed thicknesses of the two layers.
//
// The contact distance (~0.7-1 mm) is about twice the crossing spacing, so the summed node-node
// penalties form a smooth sheet-like barrier that still opens where material is missing. Detection
// runs once per substep on a spatial hash keyed by (cell, layer); it keeps the K nearest foreign
// nodes within reach and a directional bound per node. The terms are re-evaluated at every Newton
// iterate from the stored candidates; each node writes only its own row of ext (no scatter).
//
// Order independence: the hash is filled with atomics, so a cell's nodes come in any order. The
// reference keeps the first of equally near candidates it meets; here the nearest K are the
// smallest (distance, node index) pairs and the nearest foreign node the smallest (clearance, node
// index) pair, so results do not depend on the fill order. They equal the reference's for the
// order that meets tied nodes in ascending index. A slot that two of the 27 neighbour cells share
// is visited twice, as in the reference, so a node can fill two of the K places.
//
// Arithmetic is FP32 with every product rounded as written (contract off, correctly rounded divide
// and sqrt): candidates, bounds and terms are bit-identical to a scalar FP32 evaluation of the same
// expressions in this order.
//
// Layout: cells 2 x 2^20 u32 (counts scanned into starts, then fill cursors); items N u32 node ids;
// pairs K planes of Np; cmeta 2 float4 per node (own bound; reciprocal bound, always unbounded
// here); ext 9 planes (gradient xyz, Hessian xx xy xz yy yz zz).
#pragma clang fp contract(off)
#include <stdexcept>

#include "bd_internal.h"
#include "bounds.h"

namespace bd {

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kMask = (uint32_t)kPlyTable - 1u;
// CONTACT (gpu/contact.ts:8-18)
constexpr float kCell = 1.0e-3f;
constexpr float kStiff = 3000.f;
constexpr float kFriction = 0.45f;
constexpr float kFrictionSpeed = 0.002f;
constexpr float kBoundFraction = 0.45f;
constexpr float kFar = 1e9f;
constexpr int kScanBlock = 1024;  // counts per scan block (256 threads x 4)

struct PlyArgs {
  int N, Np, nl;
  const float* x;
  const float* dx;
  const uint32_t* ijl;
  uint32_t* cells;
  uint32_t* items;
  uint32_t* pairs;
  float4* cmeta;
  float* ext;
  float d01, d02, d12;
  float eps_u;  // friction regularisation: speed x h
};

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float len(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 cur(const PlyArgs& A, uint32_t n) {
  const int Np = A.Np;
  return {A.x[n] + A.dx[n], A.x[Np + n] + A.dx[Np + n], A.x[2 * Np + n] + A.dx[2 * Np + n]};
}
__device__ __forceinline__ V3 step(const PlyArgs& A, uint32_t n) {
  return {A.dx[n], A.dx[A.Np + n], A.dx[2 * A.Np + n]};
}
__device__ __forceinline__ uint32_t layer_of(const PlyArgs& A, uint32_t n) { return A.ijl[n] >> 24; }

// cell index of a coordinate; clamped so that far or non-finite positions stay well defined
__device__ __forceinline__ int cell_of(float p) {
  float q = floorf(p / kCell);
  return (int)fminf(fmaxf(q, -1e9f), 1e9f);
}
__device__ __forceinline__ uint32_t hash_cell(int cx, int cy, int cz, uint32_t layer) {
  return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u) ^
          (layer * 2654435761u)) & kMask;
}
__device__ __forceinline__ uint32_t node_hash(const PlyArgs& A, uint32_t n) {
  V3 p = cur(A, n);
  return hash_cell(cell_of(p.x), cell_of(p.y), cell_of(p.z), layer_of(A, n));
}

// contact distance of a layer pair (contact.wgsl:51-57)
__device__ __forceinline__ float pair_distance(const PlyArgs& A, uint32_t la, uint32_t lb) {
  const uint32_t s = la + lb;
  return s == 1u ? A.d01 : (s == 2u ? A.d02 : A.d12);
}

// ---------------------------------------------------------------- hash

__global__ __launch_bounds__(256) void k_ply_clear(uint32_t* cells) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 2 * kPlyTable) cells[t] = 0u;
}

__global__ __launch_bounds__(256) void k_ply_count(PlyArgs A) {
  uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N) return;
  atomicAdd(&A.cells[node_hash(A, n)], 1u);
}

// exclusive scan of 1024 counts per block; the block total goes to sums[block]
__global__ __launch_bounds__(256) void k_ply_scan_blocks(uint32_t* data, uint32_t* sums) {
  __shared__ uint32_t sh[256];
  const int t = threadIdx.x;
  const size_t base = (size_t)blockIdx.x * kScanBlock + 4 * t;
  uint32_t v[4];
  for (int q = 0; q < 4; q++) v[q] = data[base + q];
  const uint32_t tot = (v[0] + v[1]) + (v[2] + v[3]);
  sh[t] = tot;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    uint32_t a = t >= off ? sh[t - off] : 0u;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  uint32_t run = sh[t] - tot;
  for (int q = 0; q < 4; q++) {
    data[base + q] = run;
    run += v[q];
  }
  if (t == 255) sums[blockIdx.x] = sh[255];
}

// adds the scanned block offsets (sums, itself scanned by one k_ply_scan_blocks block)
__global__ __launch_bounds__(256) void k_ply_scan_add(uint32_t* data, const uint32_t* sums) {
  const uint32_t add = sums[blockIdx.x];
  const size_t base = (size_t)blockIdx.x * kScanBlock;
  for (int q = threadIdx.x; q < kScanBlock; q += 256) data[base + q] += add;
}

__global__ __launch_bounds__(256) void k_ply_fill(PlyArgs A) {
  uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N) return;
  const uint32_t h = node_hash(A, n);
  const uint32_t slot = A.cells[h] + atomicAdd(&A.cells[kPlyTable + h], 1u);
  A.items[slot] = n;
}

// ---------------------------------------------------------------- detection (contact.wgsl:83-137)

__global__ __launch_bounds__(256) void k_ply_detect(PlyArgs A) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N) return;
  const V3 p = cur(A, n);
  const uint32_t ln = layer_of(A, n);
  uint32_t best[kContactK];
  float bestD[kContactK];
  for (int k = 0; k < kContactK; k++) {
    best[k] = kNone;
    bestD[k] = kFar;
  }
  float dmin = kFar;
  uint32_t mmin = kNone;
  V3 nmin = {0.f, 0.f, 0.f};
  const int cx = cell_of(p.x), cy = cell_of(p.y), cz = cell_of(p.z);
  const float reach = 0.25f * kCell;
  for (uint32_t lo = 0; lo < (uint32_t)A.nl; lo++) {
    if (lo == ln) continue;
    const float r = pair_distance(A, ln, lo);
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dxc = -1; dxc <= 1; dxc++) {
          const uint32_t h = hash_cell(cx + dxc, cy + dy, cz + dz, lo);
          const uint32_t start = A.cells[h], count = A.cells[kPlyTable + h];
          for (uint32_t q = 0; q < count; q++) {
            const uint32_t m = A.items[start + q];
            if (layer_of(A, m) != lo) continue;
            const V3 dv = sub(p, cur(A, m));
            const float d = len(dv);
            const float e = d - r;
            if ((e < dmin || (e == dmin && m < mmin)) && d > 1e-9f) {
              dmin = e;
              mmin = m;
              nmin = {dv.x / d, dv.y / d, dv.z / d};
            }
            if (d >= r + reach) continue;
            uint32_t cand = m;
            float cd = e;
            for (int k = 0; k < kContactK; k++) {
              if (cd < bestD[k] || (cd == bestD[k] && cand < best[k])) {
                const uint32_t ts = best[k];
                const float td = bestD[k];
                best[k] = cand;
                bestD[k] = cd;
                cand = ts;
                cd = td;
              }
            }
          }
        }
  }
  for (int k = 0; k < kContactK; k++) A.pairs[(size_t)k * A.Np + n] = best[k];
  // directional bound (newtonApply): the step may approach the nearest foreign node only by a share
  // of its clearance plus half the contact distance of layers 0 and 1
  A.cmeta[2 * (size_t)n] = dmin < 1e8f ? make_float4(nmin.x, nmin.y, nmin.z, kBoundFraction * (fmaxf(0.f, dmin) + 0.5f * A.d01))
                                       : make_float4(0.f, 0.f, 0.f, kUnbounded);
  A.cmeta[2 * (size_t)n + 1] = make_float4(0.f, 0.f, 0.f, kUnbounded);
}

// ---------------------------------------------------------------- terms (contact.wgsl:139-164)

__global__ __launch_bounds__(256) void k_ply_terms(PlyArgs A) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N) return;
  float g[3] = {0.f, 0.f, 0.f};
  float H[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};  // H[c][r]: column c, row r
  const V3 p = cur(A, n);
  const V3 sn = step(A, n);
  const uint32_t ln = layer_of(A, n);
  for (int k = 0; k < kContactK; k++) {
    const uint32_t m = A.pairs[(size_t)k * A.Np + n];
    if (m == kNone) continue;
    const V3 dv = sub(p, cur(A, m));
    const float d = len(dv);
    const float r = pair_distance(A, ln, layer_of(A, m));
    if (d >= r || d <= 1e-9f) continue;
    const float nn[3] = {dv.x / d, dv.y / d, dv.z / d};
    const float lam = kStiff * (r - d);
    for (int a = 0; a < 3; a++) g[a] = g[a] - lam * nn[a];
    float nnT[3][3];
    for (int c = 0; c < 3; c++)
      for (int rr = 0; rr < 3; rr++) nnT[c][rr] = nn[rr] * nn[c];
    for (int c = 0; c < 3; c++)
      for (int rr = 0; rr < 3; rr++) H[c][rr] = H[c][rr] + kStiff * nnT[c][rr];
    // lagged-normal regularised Coulomb friction on the relative tangential step
    const V3 rel = sub(sn, step(A, m));
    const float rn = (rel.x * nn[0] + rel.y * nn[1]) + rel.z * nn[2];
    const V3 ut = {rel.x - rn * nn[0], rel.y - rn * nn[1], rel.z - rn * nn[2]};
    const float kf = kFriction * lam / fmaxf(len(ut), A.eps_u);
    g[0] = g[0] + kf * ut.x;
    g[1] = g[1] + kf * ut.y;
    g[2] = g[2] + kf * ut.z;
    for (int c = 0; c < 3; c++)
      for (int rr = 0; rr < 3; rr++) H[c][rr] = H[c][rr] + kf * ((rr == c ? 1.f : 0.f) - nnT[c][rr]);
  }
  const size_t Np = (size_t)A.Np;
  A.ext[n] = g[0];
  A.ext[Np + n] = g[1];
  A.ext[2 * Np + n] = g[2];
  A.ext[3 * Np + n] = H[0][0];
  A.ext[4 * Np + n] = H[0][1];
  A.ext[5 * Np + n] = H[0][2];
  A.ext[6 * Np + n] = H[1][1];
  A.ext[7 * Np + n] = H[1][2];
  A.ext[8 * Np + n] = H[2][2];
}

PlyArgs make_args(Solver& s) {
  PlyArgs A;
  A.N = s.N;
  A.Np = s.Np;
  A.nl = s.nl;
  A.x = s.x;
  A.dx = s.dx;
  A.ijl = s.ijl;
  A.cells = s.pcells;
  A.items = s.pitems;
  A.pairs = s.pairs;
  A.cmeta = s.cmeta;
  A.ext = s.ext;
  A.d01 = s.ply_dist[0];
  A.d02 = s.ply_dist[1];
  A.d12 = s.ply_dist[2];
  A.eps_u = kFrictionSpeed * (float)s.P.h;
  return A;
}

}  // namespace

void ply_contact_alloc(Solver& s) {
  const size_t Np = (size_t)s.Np, N = (size_t)s.N;
  // pair distances (gpu/contact.ts:52-59): half the summed thicknesses, a missing layer counting 0
  double t[kMaxPlies] = {0.0, 0.0, 0.0};
  for (int l = 0; l < s.nl && l < kMaxPlies; l++) t[l] = s.layers_h[l].thickness;
  s.ply_dist[0] = (float)(0.5 * (t[0] + t[1]));
  s.ply_dist[1] = (float)(0.5 * (t[0] + t[2]));
  s.ply_dist[2] = (float)(0.5 * (t[1] + t[2]));
  BD_HIP(hipMalloc(&s.pcells, 2 * (size_t)kPlyTable * sizeof(uint32_t)));
  BD_HIP(hipMalloc(&s.psums, (kPlyTable / kScanBlock + 1) * sizeof(uint32_t)));
  BD_HIP(hipMalloc(&s.pitems, (N + 1) * sizeof(uint32_t)));
  BD_HIP(hipMalloc(&s.pairs, kContactK * Np * sizeof(uint32_t)));
  BD_HIP(hipMemsetAsync(s.pcells, 0, 2 * (size_t)kPlyTable * sizeof(uint32_t), s.stream));
  BD_HIP(hipMemsetAsync(s.pitems, 0, (N + 1) * sizeof(uint32_t), s.stream));
  BD_HIP(hipMemsetAsync(s.pairs, 0xFF, kContactK * Np * sizeof(uint32_t), s.stream));
}

void launch_ply_detect(Solver& s) {
  if (!s.pcells) throw std::runtime_error("cross-layer contact buffers were not created (one-layer scene)");
  static_assert(kPlyTable / kScanBlock <= kScanBlock, "the block sums are scanned by one block");
  PlyArgs A = make_args(s);
  const int g = grid_for(s.N, 256);
  {
    // clear 8 MB; count: x, dx, ijl (28 B) + an atomic; scan: 2 x 4 MB; fill: 28 B + atomic + 4 B
    ProfScope ps(s, PC_CONTACT_HASH, 2.0 * kPlyTable * 4 + s.N * 28.0 + 2.0 * kPlyTable * 4 + s.N * 36.0);
    k_ply_clear<<<grid_for(2 * kPlyTable, 256), 256, 0, s.stream>>>(s.pcells);
    k_ply_count<<<g, 256, 0, s.stream>>>(A);
    k_ply_scan_blocks<<<kPlyTable / kScanBlock, 256, 0, s.stream>>>(s.pcells, s.psums);
    k_ply_scan_blocks<<<1, 256, 0, s.stream>>>(s.psums, s.psums + kPlyTable / kScanBlock);
    k_ply_scan_add<<<kPlyTable / kScanBlock, 256, 0, s.stream>>>(s.pcells, s.psums);
    k_ply_fill<<<g, 256, 0, s.stream>>>(A);
  }
  {
    // per node: x, dx, ijl, K pairs, cmeta (minimal unique bytes)
    ProfScope ps(s, PC_CONTACT_DETECT, s.N * (28.0 + 4.0 * kContactK + 32.0));
    k_ply_detect<<<g, 256, 0, s.stream>>>(A);
  }
  s.contact_detected = true;
}

void launch_ply_terms(Solver& s) {
  if (!s.pcells) throw std::runtime_error("cross-layer contact buffers were not created (one-layer scene)");
  PlyArgs A = make_args(s);
  // per node: x, dx, ijl, K pairs; ext 36 B
  ProfScope ps(s, PC_CONTACT_TERMS, s.N * (28.0 + 4.0 * kContactK + 36.0));
  k_ply_terms<<<grid_for(s.N, 256), 256, 0, s.stream>>>(A);
}

}  // namespace bd

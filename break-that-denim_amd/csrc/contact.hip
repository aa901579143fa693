// Self-contact of a cloth layer (node against an intact lattice cell) and the floor: the spatial
// face hash, candidate detection with directional bounds, the penalty/friction terms and the
// feasibility projection of a step. Restates src/simulation/gpu/wgsl/selfcontact.wgsl:1-564,
// scan.wgsl:1-61 and step.wgsl:61-95 (projectStart, feasible, floorBound) with the constants of
// gpu/selfcontact.ts:7-33 (SELF_CONTACT) and src/simulation/floor.ts (FLOOR_CONTACT).
//
// Arithmetic is FP32 with every product rounded as written (contract off, correctly rounded
// divide and sqrt): candidate sets, bounds and the fixed-point sums come out bit-identical to a
// scalar FP32 evaluation of the same expressions. Scatter sums are fixed-point integer atomics
// (gradient 1e8 per N, Hessian 1e3 per N/m), so their order does not matter.
//
// Layout (Np-padded planes unless stated): ijl per node; cells 2 x 2^17 u32 (starts, then fill
// counts); items 2 float4 per record, <= 2 records per face: (centroid now, bounding radius),
// (centroid at the substep start, face id bits); pairs K planes; keys and marks one plane each;
// cmeta 2 float4 per node (own bound, reciprocal bound; zero normal = unbounded); acc and ext 9
// planes (gradient xyz, Hessian xx xy xz yy yz zz).
#pragma clang fp contract(off)
#include <stdexcept>

#include <climits>

#include "bd_internal.h"
#include "bounds.h"

namespace bd {

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kMask = (uint32_t)kTable - 1u;
// SELF_CONTACT (gpu/selfcontact.ts:7-33)
constexpr float kCell = 12e-3f;
constexpr float kReach = 3e-3f;
constexpr float kDelta = 1.0e-3f;
constexpr float kStiff = 500.f;
constexpr float kFriction = 0.45f;
constexpr float kFrictionSpeed = 0.002f;
constexpr float kBoundFraction = 0.45f;
// selfcontact.wgsl:22-30, 293
constexpr float kHardCore = 0.3e-3f;
constexpr float kGS = 1e8f;
constexpr float kHS = 1e3f;
constexpr float kRecipUnit = 1e-5f;
constexpr int kWindow = 3;

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 divs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float len(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 vmin(V3 a, V3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
__device__ __forceinline__ V3 vmax(V3 a, V3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
__device__ __forceinline__ V3 ld(const float* p, int Np, uint32_t n) { return {p[n], p[Np + n], p[2 * Np + n]}; }

struct CtArgs {
  int N, Np;
  Geo geo;
  const DevLayer* layers;
  const float* x;
  float* dx;
  const uint32_t* ijl;
  const uint8_t* pinned;
  const uint16_t* flags;
  const float* mass;
  uint32_t* cells;
  float4* items;
  uint32_t* pairs;
  unsigned long long* keys;
  uint32_t* marks;
  float4* cmeta;
  int* acc;
  float* ext;
  float h, gnorm;
  int floor_on;
  float floor_z;
};

__device__ __forceinline__ bool seg_active(const CtArgs& A, uint32_t s) {
  return (A.flags[s] & (SEG_EXISTS | SEG_BROKEN)) == SEG_EXISTS;
}
__device__ __forceinline__ V3 cur(const CtArgs& A, uint32_t n) { return ld(A.x, A.Np, n) + ld(A.dx, A.Np, n); }
__device__ __forceinline__ V3 x0(const CtArgs& A, uint32_t n) { return ld(A.x, A.Np, n); }

__device__ __forceinline__ uint32_t hash_cell(int cx, int cy, int cz) {
  return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u)) & kMask;
}
__device__ __forceinline__ int cell_of(float p) { return (int)floorf(p / kCell); }

struct Corners {
  uint32_t c[4];  // c00, c10, c01, c11
};
__device__ __forceinline__ Corners corners(const CtArgs& A, uint32_t f) {
  uint32_t nx = (uint32_t)A.geo.nx[A.ijl[f] >> 24];
  return {{f, f + 1u, f + nx, f + nx + 1u}};
}

// intact cell: all four spans hold (selfcontact.wgsl:78-88); every node of this build is in the solve
__device__ bool face_valid(const CtArgs& A, uint32_t f) {
  uint32_t q = A.ijl[f], l = q >> 24;
  uint32_t i = q & 0xFFFu, j = (q >> 12) & 0xFFFu;
  if (i + 1u >= (uint32_t)A.geo.nx[l] || j + 1u >= (uint32_t)A.geo.ny[l]) return false;
  Corners c = corners(A, f);
  return seg_active(A, 2u * c.c[0]) && seg_active(A, 2u * c.c[0] + 1u) && seg_active(A, 2u * c.c[1] + 1u) &&
         seg_active(A, 2u * c.c[2]);
}

__device__ __forceinline__ V3 quarter_sum(V3 a, V3 b, V3 c, V3 d) { return 0.25f * (((a + b) + c) + d); }

// hashed at the predicted centroid and, when that is another cell, at the start centroid
__device__ void hash_face(const CtArgs& A, uint32_t f, uint32_t& h1, uint32_t& h0) {
  Corners c = corners(A, f);
  V3 m1 = quarter_sum(cur(A, c.c[0]), cur(A, c.c[1]), cur(A, c.c[2]), cur(A, c.c[3]));
  V3 m0 = quarter_sum(x0(A, c.c[0]), x0(A, c.c[1]), x0(A, c.c[2]), x0(A, c.c[3]));
  int a1 = cell_of(m1.x), b1 = cell_of(m1.y), e1 = cell_of(m1.z);
  int a0 = cell_of(m0.x), b0 = cell_of(m0.y), e0 = cell_of(m0.z);
  h1 = hash_cell(a1, b1, e1);
  h0 = (a0 == a1 && b0 == b1 && e0 == e1) ? kNone : hash_cell(a0, b0, e0);
}

// closest point of triangle (a, b, c) to p: weights and squared distance (selfcontact.wgsl:92-134)
__device__ void closest_tri(V3 p, V3 a, V3 b, V3 c, float w[3], float& d2) {
  V3 ab = b - a, ac = c - a, ap = p - a;
  float d1 = dot(ab, ap), d2_ = dot(ac, ap);
  w[0] = 1.f;
  w[1] = 0.f;
  w[2] = 0.f;
  if (!(d1 <= 0.f && d2_ <= 0.f)) {
    V3 bp = p - b;
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    V3 cp = p - c;
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    float vc = d1 * d4 - d3 * d2_;
    float vb = d5 * d2_ - d1 * d6;
    float va = d3 * d6 - d5 * d4;
    if (d3 >= 0.f && d4 <= d3) {
      w[0] = 0.f;
      w[1] = 1.f;
    } else if (d6 >= 0.f && d5 <= d6) {
      w[0] = 0.f;
      w[2] = 1.f;
    } else if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
      float v = d1 / (d1 - d3);
      w[0] = 1.f - v;
      w[1] = v;
    } else if (vb <= 0.f && d2_ >= 0.f && d6 <= 0.f) {
      float v = d2_ / (d2_ - d6);
      w[0] = 1.f - v;
      w[2] = v;
    } else if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
      float v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      w[0] = 0.f;
      w[1] = 1.f - v;
      w[2] = v;
    } else {
      float den = 1.f / ((va + vb) + vc);
      float wb = vb * den, wc = vc * den;
      w[0] = (1.f - wb) - wc;
      w[1] = wb;
      w[2] = wc;
    }
  }
  V3 q = (w[0] * a + w[1] * b) + w[2] * c;
  V3 e = p - q;
  d2 = dot(e, e);
}

struct FaceHit {
  float w[4];  // c00, c10, c01, c11
  V3 q;
  float d;
};

__device__ FaceHit face_hit_p(V3 p, V3 p00, V3 p10, V3 p01, V3 p11) {
  float t1[3], t2[3], s1, s2;
  closest_tri(p, p00, p10, p11, t1, s1);
  closest_tri(p, p00, p11, p01, t2, s2);
  FaceHit o;
  if (s1 <= s2) {
    o.w[0] = t1[0];
    o.w[1] = t1[1];
    o.w[2] = 0.f;
    o.w[3] = t1[2];
    o.d = sqrtf(s1);
  } else {
    o.w[0] = t2[0];
    o.w[1] = 0.f;
    o.w[2] = t2[2];
    o.w[3] = t2[1];
    o.d = sqrtf(s2);
  }
  o.q = ((o.w[0] * p00 + o.w[1] * p10) + o.w[2] * p01) + o.w[3] * p11;
  return o;
}
__device__ __forceinline__ FaceHit face_hit(const CtArgs& A, V3 p, const Corners& c) {
  return face_hit_p(p, cur(A, c.c[0]), cur(A, c.c[1]), cur(A, c.c[2]), cur(A, c.c[3]));
}
__device__ __forceinline__ FaceHit face_hit_at(const CtArgs& A, V3 p, const Corners& c) {
  return face_hit_p(p, x0(A, c.c[0]), x0(A, c.c[1]), x0(A, c.c[2]), x0(A, c.c[3]));
}

// distance from q to segment (a, b)
__device__ __forceinline__ float seg_dist(V3 q, V3 a, V3 b) {
  V3 e = b - a;
  float s = fminf(fmaxf(dot(q - a, e) / fmaxf(dot(e, e), 1e-20f), 0.f), 1.f);
  return len((q - a) - s * e);
}

// segment (o, o + d) against triangle (a, b, c): parameter of the crossing, or 2
__device__ float seg_tri(V3 o, V3 d, V3 a, V3 b, V3 c) {
  V3 e1 = b - a, e2 = c - a;
  V3 p = cross(d, e2);
  float det = dot(e1, p);
  if (fabsf(det) < 1e-18f) return 2.f;
  float inv = 1.f / det;
  V3 s = o - a;
  float u = dot(s, p) * inv;
  if (u < 0.f || u > 1.f) return 2.f;
  V3 q = cross(s, e1);
  float v = dot(d, q) * inv;
  if (v < 0.f || u + v > 1.f) return 2.f;
  float t = dot(e2, q) * inv;
  return (t >= 0.f && t <= 1.f) ? t : 2.f;
}

__device__ float path_crossing(const CtArgs& A, V3 o, V3 d, const Corners& c) {
  V3 a0 = x0(A, c.c[0]), b0 = x0(A, c.c[1]), e0 = x0(A, c.c[2]), g0 = x0(A, c.c[3]);
  float t = fminf(seg_tri(o, d, a0, b0, g0), seg_tri(o, d, a0, g0, e0));
  V3 a1 = cur(A, c.c[0]), b1 = cur(A, c.c[1]), e1 = cur(A, c.c[2]), g1 = cur(A, c.c[3]);
  return fminf(t, fminf(seg_tri(o, d, a1, b1, g1), seg_tri(o, d, a1, g1, e1)));
}

__device__ __forceinline__ uint32_t lattice_node(int off, int nx, int i, int j) { return (uint32_t)(off + j * nx + i); }

// surviving spans along a row from i0 to i1 / along a column from j0 to j1 (selfcontact.wgsl:255-269)
__device__ bool row_holds(const CtArgs& A, int off, int nx, int j, int i0, int i1) {
  for (int i = min(i0, i1); i < max(i0, i1); i++)
    if (!seg_active(A, 2u * lattice_node(off, nx, i, j))) return false;
  return true;
}
__device__ bool column_holds(const CtArgs& A, int off, int nx, int i, int j0, int j1) {
  for (int j = min(j0, j1); j < max(j0, j1); j++)
    if (!seg_active(A, 2u * lattice_node(off, nx, i, j) + 1u)) return false;
  return true;
}
// the node still reaches the face's nearest corner along either L-shaped lattice path
__device__ bool lattice_linked(const CtArgs& A, int off, int nx, int ni, int nj, int fi, int fj) {
  int ci = min(max(ni, fi), fi + 1);
  int cj = min(max(nj, fj), fj + 1);
  if (row_holds(A, off, nx, nj, ni, ci) && column_holds(A, off, nx, ci, nj, cj)) return true;
  return column_holds(A, off, nx, ni, nj, cj) && row_holds(A, off, nx, cj, ni, ci);
}

// ---------------------------------------------------------------- hash

__global__ __launch_bounds__(256) void k_clear_cells(uint32_t* cells) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 2 * kTable) cells[t] = 0u;
}

__global__ __launch_bounds__(256) void k_face_count(CtArgs A) {
  uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= (uint32_t)A.N || !face_valid(A, f)) return;
  uint32_t h1, h0;
  hash_face(A, f, h1, h0);
  atomicAdd(&A.cells[h1], 1u);
  if (h0 != kNone) atomicAdd(&A.cells[h0], 1u);
}

// exclusive scan of the 2^17 counts: 128 blocks of 1024, then the block offsets
__global__ __launch_bounds__(256) void k_scan_blocks(uint32_t* data, uint32_t* sums) {
  __shared__ uint32_t sh[256];
  const int t = threadIdx.x;
  const int base = blockIdx.x * 1024 + 4 * t;
  uint32_t v[4];
  for (int q = 0; q < 4; q++) v[q] = data[base + q];
  uint32_t tot = v[0] + v[1] + v[2] + v[3];
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

__global__ __launch_bounds__(128) void k_scan_add(uint32_t* data, const uint32_t* sums) {
  __shared__ uint32_t sh[128];
  const int t = threadIdx.x, b = blockIdx.x;
  sh[t] = t < b ? sums[t] : 0u;
  __syncthreads();
  for (int s = 64; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  uint32_t add = sh[0];
  for (int q = t; q < 1024; q += 128) data[b * 1024 + q] += add;
}

__global__ __launch_bounds__(256) void k_face_fill(CtArgs A) {
  uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= (uint32_t)A.N || !face_valid(A, f)) return;
  Corners c = corners(A, f);
  V3 a0 = x0(A, c.c[0]), b0 = x0(A, c.c[1]), e0 = x0(A, c.c[2]), g0 = x0(A, c.c[3]);
  V3 a1 = a0 + ld(A.dx, A.Np, c.c[0]), b1 = b0 + ld(A.dx, A.Np, c.c[1]);
  V3 e1 = e0 + ld(A.dx, A.Np, c.c[2]), g1 = g0 + ld(A.dx, A.Np, c.c[3]);
  V3 m0 = quarter_sum(a0, b0, e0, g0);
  V3 m1 = quarter_sum(a1, b1, e1, g1);
  float r0 = fmaxf(fmaxf(dot(a0 - m0, a0 - m0), dot(b0 - m0, b0 - m0)), fmaxf(dot(e0 - m0, e0 - m0), dot(g0 - m0, g0 - m0)));
  float r1 = fmaxf(fmaxf(dot(a1 - m1, a1 - m1), dot(b1 - m1, b1 - m1)), fmaxf(dot(e1 - m1, e1 - m1), dot(g1 - m1, g1 - m1)));
  float4 i1 = make_float4(m1.x, m1.y, m1.z, sqrtf(fmaxf(r0, r1)));
  float4 i0 = make_float4(m0.x, m0.y, m0.z, __uint_as_float(f));
  uint32_t h1, h0;
  hash_face(A, f, h1, h0);
  uint32_t slot = A.cells[h1] + atomicAdd(&A.cells[kTable + h1], 1u);
  A.items[2 * slot] = i1;
  A.items[2 * slot + 1] = i0;
  if (h0 != kNone) {
    uint32_t s0 = A.cells[h0] + atomicAdd(&A.cells[kTable + h0], 1u);
    A.items[2 * s0] = i1;
    A.items[2 * s0 + 1] = i0;
  }
}

// ---------------------------------------------------------------- detection

// per node: a broken span within the window along its lattice row (selfcontact.wgsl:295-313)
__global__ __launch_bounds__(256) void k_damage_row(CtArgs A) {
  uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N) return;
  uint32_t q = A.ijl[n], l = q >> 24;
  int i = (int)(q & 0xFFFu), j = (int)((q >> 12) & 0xFFFu);
  int off = A.geo.off[l], nx = A.geo.nx[l];
  uint32_t mark = 0u;
  for (int a = max(i - kWindow, 0); a <= min(i + kWindow, nx - 1); a++) {
    uint32_t m = lattice_node(off, nx, a, j);
    if ((A.flags[2u * m] & SEG_BROKEN) || (A.flags[2u * m + 1u] & SEG_BROKEN)) mark = 1u;
  }
  A.marks[n] = mark;
}

// One node's detection (selfcontact.wgsl:317-457): the start, every record of the scanned cells in
// order, and the end (K candidates, the bound, the reciprocal keys of the bounding face's corners).
struct Detect {
  uint32_t n;
  bool live;  // not pinned: scans and takes a bound
  V3 p, p0, path, mid;
  float half;  // half the path's length
  bool moving, damaged;
  bool one_layer;  // every face is in the node's layer: its lattice position follows from its id
  float inv_nx;
  int lox, loy, loz, ncx, ncy, ncz;
  uint32_t nl;
  int ni, nj, off, nx;
  uint32_t best[kContactK];
  float bestD[kContactK];
  float d0min;
  V3 n0;
  uint32_t f0;
  float tc;
  uint32_t fc;
};

constexpr float kStartCap = 0.5f * kCell;

__device__ void detect_begin(const CtArgs& A, uint32_t n, Detect& D) {
  D.n = n;
  const int Np = A.Np;
  for (int k = 0; k < kContactK; k++) A.pairs[k * Np + n] = kNone;
  D.live = !A.pinned[n];
  if (!D.live) {
    A.cmeta[2 * n] = make_float4(0.f, 0.f, 0.f, kUnbounded);
    return;
  }
  D.p = cur(A, n);
  D.p0 = x0(A, n);
  D.path = D.p - D.p0;
  D.moving = dot(D.path, D.path) > kHardCore * kHardCore;
  D.one_layer = A.geo.nl == 1;
  D.mid = 0.5f * (D.p0 + D.p);
  D.half = 0.5f * len(D.path);
#pragma unroll
  for (int k = 0; k < kContactK; k++) {
    D.best[k] = kNone;
    D.bestD[k] = 1e9f;
  }
  D.d0min = 1e9f;
  D.n0 = {0.f, 0.f, 0.f};
  D.f0 = kNone;
  D.tc = 2.f;
  D.fc = kNone;
  const V3 lo3 = vmin(D.p, D.p0), hi3 = vmax(D.p, D.p0);
  // the scanned box: lo .. min(hi, lo + 3) per axis, as counts so no cell index can overflow
  D.lox = (int)floorf(lo3.x / kCell - 0.5f);
  D.loy = (int)floorf(lo3.y / kCell - 0.5f);
  D.loz = (int)floorf(lo3.z / kCell - 0.5f);
  D.ncx = (int)min((long long)(int)floorf(hi3.x / kCell + 0.5f) - D.lox, 3LL);
  D.ncy = (int)min((long long)(int)floorf(hi3.y / kCell + 0.5f) - D.loy, 3LL);
  D.ncz = (int)min((long long)(int)floorf(hi3.z / kCell + 0.5f) - D.loz, 3LL);
  const uint32_t q = A.ijl[n];
  D.nl = q >> 24;
  D.ni = (int)(q & 0xFFFu);
  D.nj = (int)((q >> 12) & 0xFFFu);
  D.off = A.geo.off[D.nl];
  D.nx = A.geo.nx[D.nl];
  D.inv_nx = 1.f / (float)D.nx;
  const int ny = A.geo.ny[D.nl];
  D.damaged = false;
  for (int b = max(D.nj - kWindow, 0); b <= min(D.nj + kWindow, ny - 1); b++)
    if (A.marks[lattice_node(D.off, D.nx, D.ni, b)] != 0u) D.damaged = true;
}

// one face record (i1: centroid now and bounding radius; i0: centroid at start and the face id)
// Pre-filter: most records of the scanned cells lie far from the node, and the tests of
// detect_record take square roots and divisions. False only where each of those tests must fail:
//  - near now / near at start: the squared distance is the very dot product that len() roots, and
//    it exceeds the squared threshold by a relative 1e-5, far beyond the rounding of the root, the
//    subtraction and the comparison (d0min only decreases, so a filter taken earlier in the node's
//    sequence is at least as permissive);
//  - on the path: the distance to the segment is at least the distance to its midpoint minus half
//    its length, here beyond the radius by 10 um, far beyond the rounding of either form;
//  - in a one-layer scene, a face of the node's exclusion window with no broken span nearby (the
//    face's lattice position is its id's, which is what the ijl record holds).
__device__ __forceinline__ bool detect_maybe(const Detect& D, float4 i1, float4 i0) {
  if (D.one_layer && !D.damaged) {
    // a face in the node's exclusion window is skipped by detect_record whatever its distance
    // row and column of the face by a float reciprocal, corrected by one either way (exact for the
    // 2^24 nodes self-contact allows)
    const int r = (int)__float_as_uint(i0.w) - D.off;
    int fj = (int)((float)r * D.inv_nx);
    fj -= fj * D.nx > r ? 1 : 0;
    fj += (fj + 1) * D.nx <= r ? 1 : 0;
    const int fi = r - fj * D.nx;
    const int di = fi - D.ni, dj = fj - D.nj;
    if (di >= -3 && di <= 2 && dj >= -3 && dj <= 2) return false;
  }
  const V3 c1 = {i1.x, i1.y, i1.z}, c0 = {i0.x, i0.y, i0.z};
  const V3 u1 = D.p - c1, u0 = D.p0 - c0;
  const float a1 = kReach + i1.w, a0 = fminf(D.d0min, kStartCap) + i1.w;
  if (dot(u1, u1) <= a1 * a1 * (1.f + 1e-5f) || dot(u0, u0) <= a0 * a0 * (1.f + 1e-5f)) return true;
  if (!D.moving) return false;
  const float a2 = i1.w + D.half + 1e-5f;
  const V3 v1 = c1 - D.mid, v0 = c0 - D.mid;
  return dot(v1, v1) <= a2 * a2 * (1.f + 1e-5f) || dot(v0, v0) <= a2 * a2 * (1.f + 1e-5f);
}

__device__ __forceinline__ void detect_record(const CtArgs& A, Detect& D, float4 i1, float4 i0) {
  if (!detect_maybe(D, i1, i0)) return;
  const V3 c1 = {i1.x, i1.y, i1.z}, c0 = {i0.x, i0.y, i0.z};
  const bool nearNow = len(D.p - c1) - i1.w < kReach;
  const bool nearStart = len(D.p0 - c0) - i1.w < fminf(D.d0min, kStartCap);
  const bool onPath = D.moving && fminf(seg_dist(c0, D.p0, D.p), seg_dist(c1, D.p0, D.p)) < i1.w;
  if (!nearNow && !nearStart && !onPath) return;
  const uint32_t f = __float_as_uint(i0.w);
  const uint32_t fq = A.ijl[f];
  const int di = (int)(fq & 0xFFFu) - D.ni, dj = (int)((fq >> 12) & 0xFFFu) - D.nj;
  if ((fq >> 24) == D.nl && di >= -3 && di <= 2 && dj >= -3 && dj <= 2 &&
      (!D.damaged || lattice_linked(A, D.off, D.nx, D.ni, D.nj, D.ni + di, D.nj + dj)))
    return;
  const Corners cf = corners(A, f);
  if (nearStart) {
    const FaceHit s0 = face_hit_at(A, D.p0, cf);
    if ((s0.d < D.d0min || (s0.d == D.d0min && f < D.f0)) && s0.d > 1e-9f) {
      D.d0min = s0.d;
      D.n0 = divs(D.p0 - s0.q, s0.d);
      D.f0 = f;
    }
  }
  if (onPath) {
    const float t = path_crossing(A, D.p0, D.path, cf);
    if (t < D.tc || (t == D.tc && t <= 1.f && f < D.fc)) {
      D.tc = t;
      D.fc = f;
    }
  }
  if (!nearNow) return;
  const FaceHit hit = face_hit(A, D.p, cf);
  if (hit.d >= kReach) return;
  uint32_t cand = f;
  float cd = hit.d;
#pragma unroll
  for (int k = 0; k < kContactK; k++) {
    if (cand == D.best[k]) break;
    if (cd < D.bestD[k] || (cd == D.bestD[k] && cand < D.best[k])) {
      const uint32_t ts = D.best[k];
      const float td = D.bestD[k];
      D.best[k] = cand;
      D.bestD[k] = cd;
      cand = ts;
      cd = td;
    }
  }
}

__device__ void detect_end(const CtArgs& A, const Detect& D) {
  const int Np = A.Np;
  const uint32_t n = D.n;
  for (int k = 0; k < kContactK; k++) A.pairs[k * Np + n] = D.best[k];
  // the bound: the plane of the first face the path crosses, else the nearest face at the start
  V3 nb = D.n0;
  float db = D.d0min;
  uint32_t fb = D.f0;
  if (D.fc != kNone && D.fc != D.f0) {
    const Corners c = corners(A, D.fc);
    const V3 a = x0(A, c.c[0]);
    V3 nrm = cross(x0(A, c.c[1]) - a, x0(A, c.c[2]) - a);
    const float ln = len(nrm);
    if (ln > 1e-12f) {
      nrm = divs(nrm, ln);
      const float side = dot(D.p0 - a, nrm);
      nb = side >= 0.f ? nrm : V3{-nrm.x, -nrm.y, -nrm.z};
      db = fabsf(side);
      fb = D.fc;
    }
  }
  if (fb == kNone || (fb == D.f0 && db >= kStartCap)) {
    A.cmeta[2 * n] = make_float4(0.f, 0.f, 0.f, kUnbounded);
    return;
  }
  const float share = kBoundFraction * fmaxf(db - kHardCore, 0.f);
  A.cmeta[2 * n] = make_float4(nb.x, nb.y, nb.z, share);
  // the tightest share wins, then the larger node index: the share code above the node, as the reference's
  // 32-bit key (selfcontact.wgsl:462-474), but with the node in a full 32-bit word so any node count fits
  const unsigned long long key =
      ((unsigned long long)(4095u - min(4095u, (uint32_t)(share / kRecipUnit))) << 32) | (unsigned long long)(n + 1u);
  const Corners c = corners(A, fb);
  for (int k = 0; k < 4; k++) atomicMax(&A.keys[c.c[k]], key);
}

// every record of the node's scanned cells, cell by cell in (z, y, x) order, each cell's records in
// storage order, read from the hash in global memory
__device__ void detect_scan_global(const CtArgs& A, Detect& D) {
  for (int dz = 0; dz <= D.ncz; dz++)
    for (int dy = 0; dy <= D.ncy; dy++)
      for (int dx_ = 0; dx_ <= D.ncx; dx_++) {
        const uint32_t h = hash_cell((int)((uint32_t)D.lox + (uint32_t)dx_), (int)((uint32_t)D.loy + (uint32_t)dy),
                                     (int)((uint32_t)D.loz + (uint32_t)dz));
        const uint32_t start = A.cells[h], count = A.cells[kTable + h];
        for (uint32_t r = 0; r < count; r++) detect_record(A, D, A.items[2 * (start + r)], A.items[2 * (start + r) + 1]);
      }
}

__global__ __launch_bounds__(256) void k_detect(CtArgs A) {
  const uint32_t n = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N) return;
  Detect D;
  detect_begin(A, n, D);
  if (!D.live) return;
  detect_scan_global(A, D);
  detect_end(A, D);
}

// The same detection for a one-layer scene by 16 x 16 node tiles: the tile's nodes scan boxes of a
// few shared cells, so each cell's records are staged in LDS once per tile (kStage at a time) and
// every node whose box holds the cell reads them from there. The cells of the union of the tile's
// boxes are visited in (z, y, x) order, so every node still sees its own cells in its own order and
// each cell's records in storage order: the same sequence of detect_record calls as k_detect, hence
// the same results bit for bit. A tile whose union box exceeds kUnionCells (fast nodes) scans from
// global memory as k_detect does.
constexpr int kTile = 16;
constexpr int kStage = 256;
constexpr int kUnionCells = 64;

__global__ __launch_bounds__(256) void k_detect_tiled(CtArgs A, int tiles_x) {
  __shared__ float4 rec[2 * kStage];
  __shared__ int box[6];
  __shared__ uint32_t span[2];
  const int tile = blockIdx.x;
  const int ti = tile % tiles_x, tj = tile / tiles_x;
  const int nx = A.geo.nx[0], ny = A.geo.ny[0];
  const int i = ti * kTile + (threadIdx.x % kTile), j = tj * kTile + (threadIdx.x / kTile);
  const bool inside = i < nx && j < ny;
  const uint32_t n = inside ? (uint32_t)(A.geo.off[0] + j * nx + i) : 0u;
  Detect D;
  D.live = false;
  if (inside) detect_begin(A, n, D);
  // the union of the live nodes' scanned boxes
  if (threadIdx.x < 6) box[threadIdx.x] = threadIdx.x < 3 ? INT_MAX : INT_MIN;
  __syncthreads();
  // a box end past INT_MAX (positions far out of range) leaves the union to the global scan
  const bool wraps = D.live && ((long long)D.lox + D.ncx > INT_MAX || (long long)D.loy + D.ncy > INT_MAX ||
                                (long long)D.loz + D.ncz > INT_MAX);
  if (D.live && !wraps) {
    atomicMin(&box[0], D.lox);
    atomicMin(&box[1], D.loy);
    atomicMin(&box[2], D.loz);
    atomicMax(&box[3], D.lox + D.ncx);
    atomicMax(&box[4], D.loy + D.ncy);
    atomicMax(&box[5], D.loz + D.ncz);
  }
  const bool any_live = __syncthreads_or(D.live);
  const bool any_wrap = __syncthreads_or(wraps);
  if (!any_live) return;  // uniform
  const int bx0 = box[0], by0 = box[1], bz0 = box[2], bx1 = box[3], by1 = box[4], bz1 = box[5];
  const long long cells = bx0 > bx1 ? 0 : ((long long)bx1 - bx0 + 1) * ((long long)by1 - by0 + 1) * ((long long)bz1 - bz0 + 1);
  if (any_wrap || cells > kUnionCells) {
    if (D.live) {
      detect_scan_global(A, D);
      detect_end(A, D);
    }
    return;
  }
  for (int cz = bz0; cz <= bz1; cz++)
    for (int cy = by0; cy <= by1; cy++)
      for (int cx = bx0; cx <= bx1; cx++) {
        const bool mine = D.live && cx >= D.lox && cx <= D.lox + D.ncx && cy >= D.loy && cy <= D.loy + D.ncy &&
                          cz >= D.loz && cz <= D.loz + D.ncz;
        if (!__syncthreads_or(mine)) continue;
        if (threadIdx.x == 0) {
          const uint32_t h = hash_cell(cx, cy, cz);
          span[0] = A.cells[h];
          span[1] = A.cells[kTable + h];
        }
        __syncthreads();
        const uint32_t start = span[0], count = span[1];
        for (uint32_t r0 = 0; r0 < count; r0 += kStage) {
          const uint32_t m = min((uint32_t)kStage, count - r0);
          for (uint32_t q = threadIdx.x; q < 2 * m; q += blockDim.x) rec[q] = A.items[2 * (start + r0) + q];
          __syncthreads();
          if (mine) {
            // the chunk's records this node may take, as bits, then those in order: the lanes of
            // a wave diverge over their own few records instead of all lanes taking every record
            // that any lane takes
            uint32_t take[kStage / 32];
#pragma unroll
            for (int w = 0; w < kStage / 32; w++) {
              uint32_t bits = 0u;
              const uint32_t r0 = 32u * w, nb = r0 < m ? min(32u, m - r0) : 0u;
#pragma unroll 4
              for (uint32_t b = 0; b < nb; b++)
                bits |= detect_maybe(D, rec[2 * (r0 + b)], rec[2 * (r0 + b) + 1]) ? 1u << b : 0u;
              take[w] = bits;
            }
#pragma unroll
            for (int w = 0; w < kStage / 32; w++)
              while (take[w]) {
                const uint32_t r = 32u * w + (uint32_t)__builtin_ctz(take[w]);
                take[w] &= take[w] - 1u;
                detect_record(A, D, rec[2 * r], rec[2 * r + 1]);
              }
          }
          __syncthreads();
        }
      }
  if (D.live) detect_end(A, D);
}

// reciprocal bounds of face corners; clears the keys (selfcontact.wgsl:470-484)
__global__ __launch_bounds__(256) void k_resolve(CtArgs A) {
  uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N) return;
  const unsigned long long key = A.keys[n];
  A.keys[n] = 0ull;
  if (key == 0ull) {
    A.cmeta[2 * n + 1] = make_float4(0.f, 0.f, 0.f, kUnbounded);
    return;
  }
  uint32_t j = (uint32_t)(key & 0xFFFFFFFFull) - 1u;
  float share = (float)(4095u - (uint32_t)(key >> 32)) * kRecipUnit;
  float4 o = A.cmeta[2 * j];
  A.cmeta[2 * n + 1] = make_float4(-o.x, -o.y, -o.z, share);
}

// ---------------------------------------------------------------- terms (every Newton iteration)

__device__ __forceinline__ int fixq(float v, float scale) { return (int)fminf(fmaxf(v * scale, -2e9f), 2e9f); }

__device__ void add_term(const CtArgs& A, uint32_t n, V3 g, const float H[6]) {
  const int Np = A.Np;
  atomicAdd(&A.acc[n], fixq(g.x, kGS));
  atomicAdd(&A.acc[Np + n], fixq(g.y, kGS));
  atomicAdd(&A.acc[2 * Np + n], fixq(g.z, kGS));
  for (int k = 0; k < 6; k++) atomicAdd(&A.acc[(3 + k) * Np + n], fixq(H[k], kHS));
}

__global__ __launch_bounds__(256) void k_terms(CtArgs A) {
  uint32_t n = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N) return;
  const int Np = A.Np;
  V3 p = cur(A, n);
  const float epsU = kFrictionSpeed * A.h;
  float4 own = A.cmeta[2 * n];
  V3 ownN = {own.x, own.y, own.z};
  float reach = 1e9f;
  float sWeight = sqrtf(((A.mass[n] * A.gnorm) * kDelta) / kStiff);
  V3 dn = ld(A.dx, Np, n);
  for (int k = 0; k < kContactK; k++) {
    uint32_t f = A.pairs[k * Np + n];
    if (f == kNone) continue;
    Corners c = corners(A, f);
    FaceHit hit = face_hit(A, p, c);
    if (hit.d <= 1e-9f) continue;
    V3 nn = divs(p - hit.q, hit.d);
    float a = dot(nn, ownN);
    if (a > 0.1f) reach = fminf(reach, ((hit.d - kDelta) + fmaxf(sWeight, 2.f * fmaxf(kDelta - hit.d, 0.f))) / a);
    if (hit.d >= kDelta) continue;
    float s = kDelta - hit.d;
    float lam = ((kStiff * s) * s) / kDelta;
    V3 rel = dn - (((hit.w[0] * ld(A.dx, Np, c.c[0]) + hit.w[1] * ld(A.dx, Np, c.c[1])) + hit.w[2] * ld(A.dx, Np, c.c[2])) +
                   hit.w[3] * ld(A.dx, Np, c.c[3]));
    V3 ut = rel - dot(rel, nn) * nn;
    float kf = (kFriction * lam) / fmaxf(len(ut), epsU);
    V3 g = (-lam) * nn + kf * ut;
    float c1 = ((2.f * kStiff) * s) / kDelta;
    float nv[3] = {nn.x, nn.y, nn.z};
    const int ri[6] = {0, 0, 0, 1, 1, 2}, ci[6] = {0, 1, 2, 1, 2, 2};
    float H[6], H2[6];
    for (int e = 0; e < 6; e++) {
      float t = nv[ri[e]] * nv[ci[e]];
      H[e] = c1 * t + kf * ((ri[e] == ci[e] ? 1.f : 0.f) - t);
      H2[e] = 2.f * H[e];
    }
    add_term(A, n, g, H2);
    for (int qn = 0; qn < 4; qn++) {
      float wq = hit.w[qn];
      if (wq <= 0.f) continue;
      float Hq[6];
      float s2 = 2.f * wq;
      for (int e = 0; e < 6; e++) Hq[e] = s2 * H[e];
      add_term(A, c.c[qn], (-wq) * g, Hq);
    }
  }
  if (reach < 1e9f && (own.x != 0.f || own.y != 0.f || own.z != 0.f)) {
    float approach = -dot(dn, ownN);
    A.cmeta[2 * n] = make_float4(own.x, own.y, own.z, fminf(own.w, approach + reach));
  }
}

__global__ __launch_bounds__(256) void k_gather(CtArgs A) {
  uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N) return;
  const int Np = A.Np;
  for (int k = 0; k < 9; k++) {
    int a = A.acc[k * Np + n];
    A.acc[k * Np + n] = 0;
    A.ext[k * Np + n] = (float)a / (k < 3 ? kGS : kHS);
  }
}

// ---------------------------------------------------------------- feasibility (step.wgsl:85-95)

__global__ __launch_bounds__(256) void k_project_start(CtArgs A) {
  uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N || A.pinned[n]) return;
  const int Np = A.Np;
  float d[3] = {A.dx[n], A.dx[Np + n], A.dx[2 * Np + n]};
  feasible_step(d, A.cmeta[2 * n], A.cmeta[2 * n + 1]);
  if (A.floor_on) {
    float half = 0.5f * (float)A.layers[A.ijl[n] >> 24].thickness;
    d[2] = floor_bound_z(A.x[2 * Np + n], 0.f, d[2], half, A.mass[n], A.gnorm, A.floor_z);
  }
  A.dx[n] = d[0];
  A.dx[Np + n] = d[1];
  A.dx[2 * Np + n] = d[2];
}

CtArgs make_args(Solver& s) {
  CtArgs A;
  A.N = s.N;
  A.Np = s.Np;
  A.geo = s.geo;
  A.layers = s.layers;
  A.x = s.x;
  A.dx = s.dx;
  A.ijl = s.ijl;
  A.pinned = s.pinned;
  A.flags = s.flags;
  A.mass = s.mass;
  A.cells = s.cells;
  A.items = s.items;
  A.pairs = s.pairs;
  A.keys = s.keys;
  A.marks = s.marks;
  A.cmeta = s.cmeta;
  A.acc = s.acc;
  A.ext = s.ext;
  A.h = (float)s.P.h;
  float gx = (float)s.P.gravity[0], gy = (float)s.P.gravity[1], gz = (float)s.P.gravity[2];
  A.gnorm = sqrtf((gx * gx + gy * gy) + gz * gz);
  A.floor_on = s.floor_on ? 1 : 0;
  A.floor_z = s.floor_z;
  return A;
}

__global__ __launch_bounds__(256) void k_reset_bounds(float4* cmeta, int n2) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n2) cmeta[t] = make_float4(0.f, 0.f, 0.f, kUnbounded);
}

}  // namespace

void contact_alloc(Solver& s) {
  const size_t Np = (size_t)s.Np, N = (size_t)s.N;
  BD_HIP(hipMalloc(&s.cells, 2 * (size_t)kTable * sizeof(uint32_t)));
  BD_HIP(hipMalloc(&s.scan_sums, 128 * sizeof(uint32_t)));
  BD_HIP(hipMalloc(&s.items, 4 * N * sizeof(float4)));
  BD_HIP(hipMalloc(&s.pairs, kContactK * Np * sizeof(uint32_t)));
  BD_HIP(hipMalloc(&s.keys, Np * sizeof(unsigned long long)));
  BD_HIP(hipMalloc(&s.marks, Np * sizeof(uint32_t)));
  BD_HIP(hipMalloc(&s.acc, 9 * Np * sizeof(int)));
  BD_HIP(hipMemsetAsync(s.cells, 0, 2 * (size_t)kTable * sizeof(uint32_t), s.stream));
  BD_HIP(hipMemsetAsync(s.pairs, 0xFF, kContactK * Np * sizeof(uint32_t), s.stream));
  BD_HIP(hipMemsetAsync(s.keys, 0, Np * sizeof(unsigned long long), s.stream));
  BD_HIP(hipMemsetAsync(s.marks, 0, Np * sizeof(uint32_t), s.stream));
  BD_HIP(hipMemsetAsync(s.acc, 0, 9 * Np * sizeof(int), s.stream));
}

void contact_reset_bounds(Solver& s) {
  k_reset_bounds<<<grid_for(2 * s.N, 256), 256, 0, s.stream>>>(s.cmeta, 2 * s.N);
}

void launch_contact_detect(Solver& s) {
  if (!s.cells) throw std::runtime_error("self-contact buffers were not created (bd_params.self_contact = 0)");
  CtArgs A = make_args(s);
  const int g = grid_for(s.N, 256);
  {
    // clear 1 MB, count: read ijl + flags + 4 corners (24 B), atomics; scan 2 x 512 KB; fill: 4 corners
    // x 24 B + 64 B of records per face
    ProfScope ps(s, PC_CONTACT_HASH, 2.0 * kTable * 4 * 2 + s.N * (4.0 + 8.0 + 4 * 24.0) + 2.0 * kTable * 4 * 2 +
                                         s.N * (4.0 + 8.0 + 4 * 24.0 + 64.0));
    k_clear_cells<<<grid_for(2 * kTable, 256), 256, 0, s.stream>>>(s.cells);
    k_face_count<<<g, 256, 0, s.stream>>>(A);
    k_scan_blocks<<<kTable / 1024, 256, 0, s.stream>>>(s.cells, s.scan_sums);
    k_scan_add<<<kTable / 1024, 128, 0, s.stream>>>(s.cells, s.scan_sums);
    k_face_fill<<<g, 256, 0, s.stream>>>(A);
  }
  {
    // per node: row marks (7 x 4 B flags), x, dx, ijl, K pairs, cmeta, key (minimal unique bytes)
    ProfScope ps(s, PC_CONTACT_DETECT, s.N * (28.0 + 4.0 + 24.0 + 4.0 + 4.0 * kContactK + 32.0 + 16.0));
    k_damage_row<<<g, 256, 0, s.stream>>>(A);
    if (s.nl == 1) {
      const int tx = (s.geo.nx[0] + kTile - 1) / kTile, ty = (s.geo.ny[0] + kTile - 1) / kTile;
      k_detect_tiled<<<tx * ty, kTile * kTile, 0, s.stream>>>(A, tx);
    } else {
      k_detect<<<g, 256, 0, s.stream>>>(A);
    }
    k_resolve<<<g, 256, 0, s.stream>>>(A);
  }
  s.contact_detected = true;
}

void launch_contact_terms(Solver& s) {
  if (!s.cells) throw std::runtime_error("self-contact buffers were not created (bd_params.self_contact = 0)");
  CtArgs A = make_args(s);
  const int g = grid_for(s.N, 256);
  // per node: x, dx, K pairs, cmeta own, mass; accumulators 36 B read+write; ext 36 B
  ProfScope ps(s, PC_CONTACT_TERMS, s.N * (24.0 + 4.0 * kContactK + 16.0 + 4.0 + 72.0 + 36.0));
  k_terms<<<g, 256, 0, s.stream>>>(A);
  k_gather<<<g, 256, 0, s.stream>>>(A);
}

void launch_project_start(Solver& s) {
  CtArgs A = make_args(s);
  ProfScope ps(s, PC_PREDICT_COMMIT, s.N * (24.0 + 32.0 + 12.0 + 4.0));
  k_project_start<<<grid_for(s.N, 256), 256, 0, s.stream>>>(A);
}

}  // namespace bd

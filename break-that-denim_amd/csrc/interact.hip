// Grips and picking: the product's input path into the solve. Restates
// src/simulation/gpu/wgsl/interact.wgsl:
//   k_grip_init      gripInit (interact.wgsl:183-196)   offsets of a grip's patch from its picked node
//                    at grab time, and the handle's origin (the picked node's position then)
//   k_grip_lead      gripLead (:80-156)  once per substep from committed positions: the pads' closing
//                    axis, the held patch's weighted centroid, entries torn out from under the pads
//                    slip out (weight 0), and the effective handle clamped to within the largest lead
//                    of the centroid (bounded total force)
//   k_grip_terms     gripTerms (:158-179) per Newton iteration: each entry's quadratic spring to its
//                    target (handle + gathered offset) into ext, added to the contact terms
//   k_pick_*         pickRay, pickRayNode, pickCandidates, pickOcclude, pickVisible, pickWrite,
//                    pickWindow, pickNear (:197-456): the nearest surface along a pointer ray (intact
//                    faces and strands over openings), else the nearest unoccluded crossing within a
//                    cursor disc, else (staged grabs) the nearest crossing with material to a point;
//                    and the connectivity window around the picked node.
// Up to kGrips grips hold disjoint patches of at most kSlotEntries entries each, so every ext entry
// has a single writer.
#include <stdexcept>

#include "bd_internal.h"

namespace bd {

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kMaxCand = BD_PICK_MAX_CANDIDATES;
constexpr float kOccludeMargin = 0.5e-3f;

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

struct IaArgs {
  int N, Np;
  Geo geo;
  const float* x;
  const float* dx;
  const uint8_t* state;
  const uint16_t* flags;
  GripSlot* slots;       // kGrips
  GripEntry* entries;    // kGrips x kSlotEntries
  float* handle;         // 12 floats per grip: effective handle, origin, closing axis (xyz + pad)
  float* ext;            // 9 planes
  int adds;              // add to ext (contact wrote it this iteration) or overwrite
  PickArgs pk;
  uint32_t* pick;        // 4 + 2 kMaxCand words
  uint32_t* pick_out;    // 8 + 256 words
};

__device__ __forceinline__ V3 pos(const IaArgs& A, uint32_t n) { return {A.x[n], A.x[A.Np + n], A.x[2 * A.Np + n]}; }
__device__ __forceinline__ V3 cur(const IaArgs& A, uint32_t n) {
  return {A.x[n] + A.dx[n], A.x[A.Np + n] + A.dx[A.Np + n], A.x[2 * A.Np + n] + A.dx[2 * A.Np + n]};
}
// a node of the cloth solve (detached and retired nodes no longer respond to a grip)
__device__ __forceinline__ bool live(const IaArgs& A, uint32_t n) { return A.state[n] == kStateCloth; }
__device__ __forceinline__ V3 ldv(const float* p) { return {p[0], p[1], p[2]}; }

// target offset of an entry at pinch scale s (interact.wgsl:74-78)
__device__ __forceinline__ V3 gathered(const GripEntry& e, const GripSlot& S, V3 axis) {
  V3 o = {e.offset[0], e.offset[1], e.offset[2]};
  if (e.side == 0.f) return S.pinch * o;
  return o - (e.side * (1.f - S.pinch) * S.pad_pitch) * axis;
}

__global__ __launch_bounds__(64) void k_grip_init(IaArgs A, int g) {
  const GripSlot& S = A.slots[g];
  const uint32_t picked = S.picked;
  const V3 pp = pos(A, picked);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    float* H = A.handle + 12 * g;
    H[0] = pp.x;
    H[1] = pp.y;
    H[2] = pp.z;
    H[4] = pp.x;
    H[5] = pp.y;
    H[6] = pp.z;
  }
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S.count) return;
  GripEntry& e = A.entries[g * kSlotEntries + k];
  const V3 o = pos(A, e.node) - pp;
  e.offset[0] = o.x;
  e.offset[1] = o.y;
  e.offset[2] = o.z;
}

// one workgroup of 256 per grip
__global__ __launch_bounds__(256) void k_grip_lead(IaArgs A) {
  __shared__ float red[4][256];
  const int g = blockIdx.x, t = threadIdx.x;
  const GripSlot& S = A.slots[g];
  const int count = S.count;
  if (count == 0) return;
  GripEntry* E = A.entries + g * kSlotEntries;
  // the closing axis: from the - pad to the + pad, as they lay at grab time
  float d[3] = {0.f, 0.f, 0.f};
  for (int k = t; k < count; k += 256)
    for (int q = 0; q < 3; q++) d[q] += E[k].side * E[k].w * E[k].offset[q];
  for (int q = 0; q < 3; q++) red[q][t] = d[q];
  __syncthreads();
  for (int r = 128; r > 0; r >>= 1) {
    if (t < r)
      for (int q = 0; q < 3; q++) red[q][t] += red[q][t + r];
    __syncthreads();
  }
  const V3 ds = {red[0][0], red[1][0], red[2][0]};
  const float ld = len(ds);
  const V3 axis = ld > 1e-12f ? (1.f / fmaxf(ld, 1e-20f)) * ds : V3{0.f, 0.f, 0.f};
  __syncthreads();
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int k = t; k < count; k += 256) {
    const GripEntry& e = E[k];
    if (!live(A, e.node)) continue;
    const V3 p = pos(A, e.node) - gathered(e, S, axis);
    acc[0] += e.w * p.x;
    acc[1] += e.w * p.y;
    acc[2] += e.w * p.z;
    acc[3] += e.w;
  }
  for (int q = 0; q < 4; q++) red[q][t] = acc[q];
  __syncthreads();
  for (int r = 128; r > 0; r >>= 1) {
    if (t < r)
      for (int q = 0; q < 4; q++) red[q][t] += red[q][t + r];
    __syncthreads();
  }
  const float hw = red[3][0];
  const V3 c = hw > 0.f ? (1.f / hw) * V3{red[0][0], red[1][0], red[2][0]} : V3{0.f, 0.f, 0.f};
  // an entry farther than a pad's radius from where the rest of the patch holds it slips out
  if (hw > 0.f && S.slip > 0.f)
    for (int k = t; k < count; k += 256) {
      GripEntry& e = E[k];
      if (e.w > 0.f && live(A, e.node) && len(pos(A, e.node) - gathered(e, S, axis) - c) > S.slip) e.w = 0.f;
    }
  if (t == 0) {
    float* H = A.handle + 12 * g;
    H[8] = axis.x;
    H[9] = axis.y;
    H[10] = axis.z;
    if (!(hw > 0.f)) {
      // nothing held is still cloth: the handle exerts nothing
      H[0] = H[4];
      H[1] = H[5];
      H[2] = H[6];
      return;
    }
    const V3 requested = V3{H[4], H[5], H[6]} + (ldv(S.goal) - ldv(S.anchor));
    V3 lead = requested - c;
    const float l = len(lead);
    if (l > S.max_lead) lead = (S.max_lead / l) * lead;
    const V3 h = c + lead;
    H[0] = h.x;
    H[1] = h.y;
    H[2] = h.z;
  }
}

__global__ __launch_bounds__(64) void k_grip_terms(IaArgs A) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = id / kSlotEntries, k = id % kSlotEntries;
  if (g >= kGrips) return;
  const GripSlot& S = A.slots[g];
  if (k >= S.count) return;
  const GripEntry& e = A.entries[id];
  if (!live(A, e.node)) return;
  const float* H = A.handle + 12 * g;
  const V3 target = V3{H[0], H[1], H[2]} + gathered(e, S, V3{H[8], H[9], H[10]});
  const V3 dd = cur(A, e.node) - target;
  const float kg = S.stiffness * e.w;
  const float keep = A.adds ? 1.f : 0.f;
  const int Np = A.Np;
  const uint32_t n = e.node;
  A.ext[n] = keep * A.ext[n] + kg * dd.x;
  A.ext[Np + n] = keep * A.ext[Np + n] + kg * dd.y;
  A.ext[2 * Np + n] = keep * A.ext[2 * Np + n] + kg * dd.z;
  A.ext[3 * Np + n] = keep * A.ext[3 * Np + n] + kg;
  A.ext[4 * Np + n] = keep * A.ext[4 * Np + n];
  A.ext[5 * Np + n] = keep * A.ext[5 * Np + n];
  A.ext[6 * Np + n] = keep * A.ext[6 * Np + n] + kg;
  A.ext[7 * Np + n] = keep * A.ext[7 * Np + n];
  A.ext[8 * Np + n] = keep * A.ext[8 * Np + n] + kg;
}

// ---------------------------------------------------------------- picking

struct Lat {
  int off, nx, ny, i, j;
};
__device__ __forceinline__ Lat lat(const IaArgs& A, uint32_t n) {
  int l = geo_layer(A.geo, (int)n);
  int r = (int)n - A.geo.off[l];
  int j = r / A.geo.nx[l];
  return {A.geo.off[l], A.geo.nx[l], A.geo.ny[l], r - j * A.geo.nx[l], j};
}
__device__ __forceinline__ bool span_on(const IaArgs& A, uint32_t owner, int f) {
  return (A.flags[2 * owner + f] & (SEG_EXISTS | SEG_BROKEN)) == SEG_EXISTS;
}
// a cloth crossing with at least one surviving incident span (interact.wgsl:206-212)
__device__ bool has_material(const IaArgs& A, uint32_t n) {
  if (!live(A, n)) return false;
  if (span_on(A, n, 0) || span_on(A, n, 1)) return true;
  Lat L = lat(A, n);
  if (L.i > 0 && span_on(A, n - 1, 0)) return true;
  return L.j > 0 && span_on(A, n - L.nx, 1);
}
// an intact cell with lower-left node f, all corners cloth (interact.wgsl:214-223)
__device__ bool face_valid(const IaArgs& A, uint32_t f) {
  Lat L = lat(A, f);
  if (L.i + 1 >= L.nx || L.j + 1 >= L.ny) return false;
  const uint32_t c10 = f + 1, c01 = f + L.nx;
  if (!live(A, f) || !live(A, c10) || !live(A, c01) || !live(A, c01 + 1)) return false;
  return span_on(A, f, 0) && span_on(A, f, 1) && span_on(A, c10, 1) && span_on(A, c01, 0);
}

// ray (o, d) against triangle (a, b, c): (t, u, v), t < 0 when missed (Moller-Trumbore)
__device__ void ray_tri(V3 o, V3 d, V3 a, V3 b, V3 c, float out[3]) {
  out[0] = -1.f;
  out[1] = out[2] = 0.f;
  const V3 e1 = b - a, e2 = c - a, p = cross(d, e2);
  const float det = dot(e1, p);
  if (fabsf(det) < 1e-14f) return;
  const float inv = 1.f / det;
  const V3 s = o - a;
  const float u = dot(s, p) * inv;
  if (u < 0.f || u > 1.f) return;
  const V3 q = cross(s, e1);
  const float v = dot(d, q) * inv;
  if (v < 0.f || u + v > 1.f) return;
  out[0] = dot(e2, q) * inv;
  out[1] = u;
  out[2] = v;
}

// ray (o, unit d) against the capsule of radius r around (a, b): (t, s), t < 0 when missed
__device__ void ray_capsule(V3 o, V3 d, V3 a, V3 b, float r, float out[2]) {
  const V3 e = b - a, w = o - a;
  const float ee = dot(e, e), de = dot(d, e), den = ee - de * de;
  float s = 0.f;
  if (den > 1e-14f) s = fminf(fmaxf((dot(e, w) * dot(d, d) - de * dot(d, w)) / den, 0.f), 1.f);
  const V3 q = a + s * e;
  const float t = fmaxf(dot(q - o, d), 0.f);
  const float gap = len(o + t * d - q);
  if (gap > r) {
    out[0] = -1.f;
    out[1] = 0.f;
    return;
  }
  out[0] = t - sqrtf(r * r - gap * gap);
  out[1] = s;
}

// nearest pointer-ray hit among the primitives node n owns (interact.wgsl:263-300)
__device__ void ray_prims(const IaArgs& A, uint32_t n, float& bt, uint32_t& bn) {
  bt = 3.4e38f;
  bn = kNone;
  const V3 o = ldv(A.pk.eye), d = ldv(A.pk.ray);
  const Lat L = lat(A, n);
  const bool fv = face_valid(A, n);
  if (fv) {
    const uint32_t c0 = n, c1 = n + 1, c2 = n + L.nx, c3 = n + L.nx + 1;
    const V3 p00 = pos(A, c0), p10 = pos(A, c1), p01 = pos(A, c2), p11 = pos(A, c3);
    float h[3];
    ray_tri(o, d, p00, p10, p11, h);
    if (h[0] > 0.f && h[0] < bt) {
      const float w0 = 1.f - h[1] - h[2], w1 = h[1], w2 = h[2];
      bt = h[0];
      bn = w0 >= fmaxf(w1, w2) ? c0 : (w1 >= w2 ? c1 : c3);
    }
    ray_tri(o, d, p00, p11, p01, h);
    if (h[0] > 0.f && h[0] < bt) {
      const float w0 = 1.f - h[1] - h[2], w1 = h[1], w2 = h[2];
      bt = h[0];
      bn = w0 >= fmaxf(w1, w2) ? c0 : (w1 >= w2 ? c3 : c2);
    }
  }
  if (!live(A, n)) return;
  for (int f = 0; f < 2; f++) {
    if (!span_on(A, n, f)) continue;
    const uint32_t other = f == 0 ? n + 1 : n + L.nx;
    if (!live(A, other)) continue;
    // the two cells beside the span: a strand covered by an intact face is picked as that face
    const bool has_side = f == 0 ? L.j > 0 : L.i > 0;
    const uint32_t side = f == 0 ? n - L.nx : n - 1;
    if (fv || (has_side && face_valid(A, side))) continue;
    float h[2];
    ray_capsule(o, d, pos(A, n), pos(A, other), A.pk.strand_radius, h);
    if (h[0] > 0.f && h[0] < bt) {
      bt = h[0];
      bn = h[1] > 0.5f ? other : n;
    }
  }
}

// projection to drawing-buffer pixels (x, y), NDC depth, clip w
__device__ void project(const IaArgs& A, V3 p, float o[4]) {
  const float* M = A.pk.view_proj;  // column-major 4 x 4
  float c[4];
  for (int r = 0; r < 4; r++) c[r] = M[r] * p.x + M[4 + r] * p.y + M[8 + r] * p.z + M[12 + r];
  const float nx = c[0] / c[3], ny = c[1] / c[3], nz = c[2] / c[3];
  o[0] = (nx * 0.5f + 0.5f) * A.pk.width;
  o[1] = (0.5f - ny * 0.5f) * A.pk.height;
  o[2] = nz;
  o[3] = c[3];
}

__global__ __launch_bounds__(256) void k_pick_ray(IaArgs A) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N) return;
  float t;
  uint32_t m;
  ray_prims(A, n, t, m);
  if (m != kNone) atomicMin(&A.pick[0], __float_as_uint(t));
}

// the material node of that hit (the lowest index on an exact tie)
__global__ __launch_bounds__(256) void k_pick_ray_node(IaArgs A) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N) return;
  const uint32_t front = __hip_atomic_load(&A.pick[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (front == kNone) return;
  float t;
  uint32_t m;
  ray_prims(A, n, t, m);
  if (m != kNone && __float_as_uint(t) == front) atomicMin(&A.pick[1], m);
}

// over no surface: crossings within the cursor disc are candidates
__global__ __launch_bounds__(256) void k_pick_candidates(IaArgs A) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N || A.pick[0] != kNone || !has_material(A, n)) return;
  float p[4];
  project(A, pos(A, n), p);
  if (p[3] <= 0.f || p[2] <= 0.f || p[2] >= 1.f) return;
  const float dxp = p[0] - A.pk.pointer[0], dyp = p[1] - A.pk.pointer[1];
  if (sqrtf(dxp * dxp + dyp * dyp) > A.pk.radius_px) return;
  const uint32_t k = atomicAdd(&A.pick[2], 1u);
  if (k < (uint32_t)kMaxCand) A.pick[4 + k] = n;
}

// a candidate is hidden when an intact face lies on the line of sight in front of it
__global__ __launch_bounds__(256) void k_pick_occlude(IaArgs A) {
  const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= (uint32_t)A.N) return;
  const uint32_t count = min(A.pick[2], (uint32_t)kMaxCand);
  if (count == 0 || !face_valid(A, f)) return;
  const Lat L = lat(A, f);
  const uint32_t c[4] = {f, f + 1, f + L.nx, f + L.nx + 1};
  const V3 p00 = pos(A, c[0]), p10 = pos(A, c[1]), p01 = pos(A, c[2]), p11 = pos(A, c[3]);
  float s[4][4];
  project(A, p00, s[0]);
  project(A, p10, s[1]);
  project(A, p01, s[2]);
  project(A, p11, s[3]);
  if (fminf(fminf(s[0][3], s[1][3]), fminf(s[2][3], s[3][3])) <= 0.f) return;
  const float R = A.pk.radius_px;
  const float lox = fminf(fminf(s[0][0], s[1][0]), fminf(s[2][0], s[3][0])) - R;
  const float loy = fminf(fminf(s[0][1], s[1][1]), fminf(s[2][1], s[3][1])) - R;
  const float hix = fmaxf(fmaxf(s[0][0], s[1][0]), fmaxf(s[2][0], s[3][0])) + R;
  const float hiy = fmaxf(fmaxf(s[0][1], s[1][1]), fmaxf(s[2][1], s[3][1])) + R;
  const float px = A.pk.pointer[0], py = A.pk.pointer[1];
  if (px < lox || py < loy || px > hix || py > hiy) return;
  const V3 o = ldv(A.pk.eye);
  for (uint32_t k = 0; k < count; k++) {
    const uint32_t m = A.pick[4 + k];
    if (m == c[0] || m == c[1] || m == c[2] || m == c[3]) continue;
    const V3 to = pos(A, m) - o;
    const float l = len(to);
    const V3 d = (1.f / l) * to;
    const float tmax = l - kOccludeMargin;
    float h1[3], h2[3];
    ray_tri(o, d, p00, p10, p11, h1);
    ray_tri(o, d, p00, p11, p01, h2);
    if ((h1[0] > 0.f && h1[0] < tmax) || (h2[0] > 0.f && h2[0] < tmax)) atomicOr(&A.pick[4 + kMaxCand + k], 1u);
  }
}

// the visible candidate nearest to the pointer on screen
__global__ __launch_bounds__(128) void k_pick_visible(IaArgs A) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= min(A.pick[2], (uint32_t)kMaxCand) || A.pick[4 + kMaxCand + k] != 0) return;
  const uint32_t n = A.pick[4 + k];
  float p[4];
  project(A, pos(A, n), p);
  const float dxp = p[0] - A.pk.pointer[0], dyp = p[1] - A.pk.pointer[1];
  const uint32_t dk = min(4095u, (uint32_t)(sqrtf(dxp * dxp + dyp * dyp) * 64.f));
  atomicMin(&A.pick[3], (dk << 20) | (n & 0xFFFFFu));
}

// the crossing with material nearest to a point within a radius (staged grabs, interact.wgsl:444-456)
__global__ __launch_bounds__(256) void k_pick_near(IaArgs A) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N || !has_material(A, n)) return;
  const float d = len(pos(A, n) - ldv(A.pk.eye));
  if (d > A.pk.strand_radius) return;
  const uint32_t dk = min(4095u, (uint32_t)(d / A.pk.strand_radius * 4095.f));
  atomicMin(&A.pick[3], (dk << 20) | (n & 0xFFFFFu));
}

__global__ void k_pick_write(IaArgs A) {
  const uint32_t front = A.pick[0];
  V3 p = {0.f, 0.f, 0.f};
  uint32_t n = kNone, kind = 0;
  if (front != kNone) {
    n = A.pick[1];
    p = ldv(A.pk.eye) + __uint_as_float(front) * ldv(A.pk.ray);
    kind = 1;
  } else if (A.pick[3] != kNone) {
    n = A.pick[3] & 0xFFFFFu;
    p = pos(A, n);
    kind = 2;
  }
  A.pick_out[0] = __float_as_uint(p.x);
  A.pick_out[1] = __float_as_uint(p.y);
  A.pick_out[2] = __float_as_uint(p.z);
  A.pick_out[3] = n;
  A.pick_out[4] = kind;
}

// connectivity around the picked node: per lattice offset in [-R, R]^2, bit 0 / 1 its own span
// along u / v survives, bit 2 not cloth, bit 3 outside the lattice (interact.wgsl:418-440)
__global__ __launch_bounds__(256) void k_pick_window(IaArgs A) {
  const int k = threadIdx.x;
  const int R = A.pk.window_radius, side = 2 * R + 1;
  if (k >= side * side) return;
  if (A.pick_out[4] == 0) {
    A.pick_out[8 + k] = 8u;
    return;
  }
  const uint32_t n = A.pick_out[3];
  const Lat L = lat(A, n);
  const int i = L.i + k % side - R, j = L.j + k / side - R;
  if (i < 0 || j < 0 || i >= L.nx || j >= L.ny) {
    A.pick_out[8 + k] = 8u;
    return;
  }
  const uint32_t m = (uint32_t)(L.off + j * L.nx + i);
  uint32_t b = 0;
  if (span_on(A, m, 0)) b |= 1u;
  if (span_on(A, m, 1)) b |= 2u;
  if (!live(A, m)) b |= 4u;
  A.pick_out[8 + k] = b;
}

IaArgs make_args(Solver& s) {
  IaArgs A;
  A.N = s.N;
  A.Np = s.Np;
  A.geo = s.geo;
  A.x = s.x;
  A.dx = s.dx;
  A.state = s.nstate;
  A.flags = s.flags;
  A.slots = s.grip_slots;
  A.entries = s.grip_entries;
  A.handle = s.grip_handle;
  A.ext = s.ext;
  A.adds = s.contact_mode != kContactNone ? 1 : 0;
  A.pk = s.pick_args;
  A.pick = s.pick;
  A.pick_out = s.pick_out;
  return A;
}

}  // namespace

void interact_alloc(Solver& s) {
  if (s.grip_slots) return;
  BD_HIP(hipMalloc(&s.grip_slots, kGrips * sizeof(GripSlot)));
  BD_HIP(hipMalloc(&s.grip_entries, (size_t)kGrips * kSlotEntries * sizeof(GripEntry)));
  BD_HIP(hipMalloc(&s.grip_handle, 12 * kGrips * sizeof(float)));
  BD_HIP(hipMalloc(&s.pick, (4 + 2 * kMaxCand) * sizeof(uint32_t)));
  BD_HIP(hipMalloc(&s.pick_out, (8 + 256) * sizeof(uint32_t)));
  BD_HIP(hipMemsetAsync(s.grip_slots, 0, kGrips * sizeof(GripSlot), s.stream));
  BD_HIP(hipMemsetAsync(s.grip_entries, 0, (size_t)kGrips * kSlotEntries * sizeof(GripEntry), s.stream));
  BD_HIP(hipMemsetAsync(s.grip_handle, 0, 12 * kGrips * sizeof(float), s.stream));
}

void launch_grip_init(Solver& s, int g) {
  IaArgs A = make_args(s);
  const int count = s.grip_slots_h[g].count;
  k_grip_init<<<grid_for(count > 0 ? count : 1, 64), 64, 0, s.stream>>>(A, g);
}

void launch_grip_lead(Solver& s) {
  if (!s.grips_active) return;
  IaArgs A = make_args(s);
  ProfScope ps(s, PC_PREDICT_COMMIT, 0.0);
  k_grip_lead<<<kGrips, 256, 0, s.stream>>>(A);
}

void launch_grip_terms(Solver& s) {
  if (!s.grips_active) return;
  IaArgs A = make_args(s);
  ProfScope ps(s, PC_CONTACT_TERMS, 0.0);
  k_grip_terms<<<kGrips * kSlotEntries / 64, 64, 0, s.stream>>>(A);
}

void launch_pick(Solver& s, bool ray) {
  IaArgs A = make_args(s);
  std::vector<uint32_t> init(4 + 2 * kMaxCand, 0u);
  init[0] = init[1] = init[3] = kNone;
  BD_HIP(hipMemcpyAsync(s.pick, init.data(), init.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s.stream));
  const int g = grid_for(s.N, 256);
  if (ray) {
    k_pick_ray<<<g, 256, 0, s.stream>>>(A);
    k_pick_ray_node<<<g, 256, 0, s.stream>>>(A);
    k_pick_candidates<<<g, 256, 0, s.stream>>>(A);
    k_pick_occlude<<<g, 256, 0, s.stream>>>(A);
    k_pick_visible<<<grid_for(kMaxCand, 128), 128, 0, s.stream>>>(A);
  } else {
    k_pick_near<<<g, 256, 0, s.stream>>>(A);
  }
  k_pick_write<<<1, 1, 0, s.stream>>>(A);
  k_pick_window<<<1, 256, 0, s.stream>>>(A);
}

}  // namespace bd

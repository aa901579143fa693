// Grips and picking: the product's input path into the solve (the reference's Interaction,
// src/simulation/gpu/interaction.ts with interact.wgsl). Same behaviour, this build's own kernels:
//
// Grips. A slot holds up to kSlotEntries patch entries (node, weight, pad side, offset from the
// grabbed node at grab time). Once per substep one 64-lane wave per slot (k_grip_wave) reduces
// with cross-lane shuffles, from committed positions: the closing axis (the weighted offsets of the
// + pad minus those of the - pad), the weighted centroid of the held material less each entry's
// pinch-gathered offset, the entries torn out from under the pads (farther than `slip` from where
// the rest of the patch holds them: weight 0), and the handle: the requested position clamped to
// within max_lead of that centroid, so the summed spring force never exceeds stiffness x weight x
// max_lead. Every Newton iteration each entry adds its spring to its target (handle + gathered
// offset) into the external terms (k_grip_springs); slots hold disjoint nodes, so each ext row has
// one writer.
//
// Picking. k_pick_surface: every node tests the primitives it owns (its cell's two triangles when
// the cell is intact, and its strands over openings as capsules) against the pointer ray and folds
// (hit distance bits, picked corner) into one 64-bit atomicMin, so the nearest hit and the lowest node
// at an exact tie come out of one pass. Without a surface hit: the crossings with material inside
// the cursor disc are compacted per wave (ballot, one atomic per wave), every intact cell whose screen
// box overlaps the disc marks the candidates it hides, and the visible candidate nearest the pointer
// on screen wins a (screen distance bits, node) atomicMin. A nearest-material pick (staged grabs) is
// the same atomicMin over (distance, node) in a sphere. Then the lattice connectivity window around
// the picked node.
#include <stdexcept>

#include "bd_internal.h"

namespace bd {

namespace {

constexpr uint32_t kNoNode = 0xFFFFFFFFu;
constexpr unsigned long long kNoKey = ~0ull;
constexpr int kCand = BD_PICK_MAX_CANDIDATES;
constexpr float kBehind = 0.5e-3f;  // an occluder must lie this far in front of the candidate

// pick scratch layout (u32 words): [0,1] surface key, [2,3] choice key, [4] candidate count,
// [5 .. 5 + kCand) candidates, [5 + kCand .. 5 + 2 kCand) hidden flags
constexpr int kPickWords = 5 + 2 * kCand;

struct F3 {
  float x, y, z;
};
__device__ __forceinline__ F3 f3(const float* p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 scale(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float norm(F3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ F3 cross3(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

struct Ctx {
  int N, Np;
  Geo geo;
  const float* x;
  const float* dx;
  const uint8_t* state;
  const uint16_t* flags;
  GripSlot* slots;
  GripEntry* entries;
  float* handle;  // per slot 12 floats: handle xyz, -, origin xyz, -, closing axis xyz, -
  float* ext;
  int accumulate;  // contact wrote ext this iteration: add to it
  PickArgs q;
  uint32_t* work;  // kPickWords
  uint32_t* out;   // 8 + 256: point xyz, node, kind, -, -, -, window
};

__device__ __forceinline__ F3 at(const Ctx& c, uint32_t n) { return {c.x[n], c.x[c.Np + n], c.x[2 * c.Np + n]}; }
__device__ __forceinline__ F3 at_iterate(const Ctx& c, uint32_t n) {
  return add(at(c, n), F3{c.dx[n], c.dx[c.Np + n], c.dx[2 * c.Np + n]});
}
__device__ __forceinline__ bool cloth(const Ctx& c, uint32_t n) { return c.state[n] == kStateCloth; }

__device__ __forceinline__ float wave_sum(float v) {
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// where an entry is held relative to the handle: its grab-time offset, pulled toward the grab line
// by the pinch (pads along the closing axis, a radial pinch by scaling)
__device__ __forceinline__ F3 held_offset(const GripEntry& e, const GripSlot& S, F3 axis) {
  const F3 o = f3(e.offset);
  if (e.side == 0.f) return scale(S.pinch, o);
  return sub(o, scale(e.side * (1.f - S.pinch) * S.pad_pitch, axis));
}

__global__ __launch_bounds__(64) void k_grip_offsets(Ctx c, int g) {
  const GripSlot& S = c.slots[g];
  const F3 p = at(c, S.picked);
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0) {
    float* H = c.handle + 12 * g;
    H[0] = H[4] = p.x;
    H[1] = H[5] = p.y;
    H[2] = H[6] = p.z;
  }
  if (k >= S.count) return;
  GripEntry& e = c.entries[g * kSlotEntries + k];
  const F3 o = sub(at(c, e.node), p);
  e.offset[0] = o.x;
  e.offset[1] = o.y;
  e.offset[2] = o.z;
}

// one wave per slot
__global__ __launch_bounds__(64) void k_grip_wave(Ctx c) {
  const int g = blockIdx.x, lane = threadIdx.x;
  const GripSlot& S = c.slots[g];
  const int count = S.count;
  if (count == 0) return;
  GripEntry* E = c.entries + g * kSlotEntries;
  float ax = 0.f, ay = 0.f, az = 0.f;
  for (int k = lane; k < count; k += 64) {
    const float sw = E[k].side * E[k].w;
    ax += sw * E[k].offset[0];
    ay += sw * E[k].offset[1];
    az += sw * E[k].offset[2];
  }
  F3 axis = {wave_sum(ax), wave_sum(ay), wave_sum(az)};
  const float la = norm(axis);
  axis = la > 1e-12f ? scale(1.f / la, axis) : F3{0.f, 0.f, 0.f};
  float sx = 0.f, sy = 0.f, sz = 0.f, sw = 0.f;
  for (int k = lane; k < count; k += 64) {
    const GripEntry& e = E[k];
    if (!cloth(c, e.node)) continue;
    const F3 h = sub(at(c, e.node), held_offset(e, S, axis));
    sx += e.w * h.x;
    sy += e.w * h.y;
    sz += e.w * h.z;
    sw += e.w;
  }
  sx = wave_sum(sx);
  sy = wave_sum(sy);
  sz = wave_sum(sz);
  sw = wave_sum(sw);
  float* H = c.handle + 12 * g;
  if (!(sw > 0.f)) {
    // no held entry is cloth any more: the handle sits at its origin and pulls nothing
    if (lane == 0) {
      H[0] = H[4];
      H[1] = H[5];
      H[2] = H[6];
      H[8] = axis.x;
      H[9] = axis.y;
      H[10] = axis.z;
    }
    return;
  }
  const F3 centre = scale(1.f / sw, F3{sx, sy, sz});
  if (S.slip > 0.f)
    for (int k = lane; k < count; k += 64) {
      GripEntry& e = E[k];
      if (e.w > 0.f && cloth(c, e.node) && norm(sub(sub(at(c, e.node), held_offset(e, S, axis)), centre)) > S.slip) e.w = 0.f;
    }
  if (lane == 0) {
    const F3 want = add(F3{H[4], H[5], H[6]}, sub(f3(S.goal), f3(S.anchor)));
    F3 lead = sub(want, centre);
    const float l = norm(lead);
    if (l > S.max_lead) lead = scale(S.max_lead / l, lead);
    const F3 h = add(centre, lead);
    H[0] = h.x;
    H[1] = h.y;
    H[2] = h.z;
    H[8] = axis.x;
    H[9] = axis.y;
    H[10] = axis.z;
  }
}

// each entry's spring at the current iterate into ext (gradient, then the isotropic Hessian k w I)
__global__ __launch_bounds__(64) void k_grip_springs(Ctx c) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = id / kSlotEntries;
  if (g >= kGrips || id % kSlotEntries >= c.slots[g].count) return;
  const GripSlot& S = c.slots[g];
  const GripEntry& e = c.entries[id];
  if (!cloth(c, e.node)) return;
  const float* H = c.handle + 12 * g;
  const F3 target = add(f3(H), held_offset(e, S, f3(H + 8)));
  const F3 d = sub(at_iterate(c, e.node), target);
  const float k = S.stiffness * e.w;
  const float terms[9] = {k * d.x, k * d.y, k * d.z, k, 0.f, 0.f, k, 0.f, k};
  for (int q = 0; q < 9; q++) {
    float* p = c.ext + (size_t)q * c.Np + e.node;
    *p = c.accumulate ? *p + terms[q] : terms[q];
  }
}

// ---------------------------------------------------------------- picking

struct Cell {
  int off, nx, ny, i, j;
};
__device__ __forceinline__ Cell cell_of(const Ctx& c, uint32_t n) {
  const int l = geo_layer(c.geo, (int)n);
  const int r = (int)n - c.geo.off[l];
  const int j = r / c.geo.nx[l];
  return {c.geo.off[l], c.geo.nx[l], c.geo.ny[l], r - j * c.geo.nx[l], j};
}
__device__ __forceinline__ bool span_alive(const Ctx& c, uint32_t owner, int fam) {
  return (c.flags[2 * owner + fam] & (SEG_EXISTS | SEG_BROKEN)) == SEG_EXISTS;
}
// a cloth crossing with at least one surviving span at it
__device__ bool has_material(const Ctx& c, uint32_t n) {
  if (!cloth(c, n)) return false;
  const Cell L = cell_of(c, n);
  return span_alive(c, n, 0) || span_alive(c, n, 1) || (L.i > 0 && span_alive(c, n - 1, 0)) ||
         (L.j > 0 && span_alive(c, n - L.nx, 1));
}
// the cell whose lower-left corner is f is intact: its four spans survive and its corners are cloth
__device__ bool cell_intact(const Ctx& c, uint32_t f) {
  const Cell L = cell_of(c, f);
  if (L.i + 1 >= L.nx || L.j + 1 >= L.ny) return false;
  const uint32_t e = f + 1, nn = f + L.nx;
  return cloth(c, f) && cloth(c, e) && cloth(c, nn) && cloth(c, nn + 1) && span_alive(c, f, 0) && span_alive(c, f, 1) &&
         span_alive(c, e, 1) && span_alive(c, nn, 0);
}

// distance along the unit ray (o, d) to triangle (a, b, c), with the barycentric weights of b and c
// (a negative distance: no hit). Solves [-d, b - a, c - a] (t, u, v) = o - a by Cramer's rule.
__device__ float ray_triangle(F3 o, F3 d, F3 a, F3 b, F3 cc, float& u, float& v) {
  const F3 e1 = sub(b, a), e2 = sub(cc, a), r = sub(o, a);
  const F3 n = cross3(e1, e2);
  const float den = -dot3(d, n);
  u = v = 0.f;
  if (fabsf(den) < 1e-14f) return -1.f;
  const float inv = 1.f / den;
  const F3 rd = cross3(r, d);
  u = dot3(e2, rd) * inv;
  v = -dot3(e1, rd) * inv;
  if (u < 0.f || v < 0.f || u + v > 1.f) return -1.f;
  return dot3(r, n) * inv;
}

// distance along the unit ray (o, d) to the capsule of radius rad around segment (a, b), with the
// segment parameter of the nearest point (a negative distance: no hit)
__device__ float ray_strand(F3 o, F3 d, F3 a, F3 b, float rad, float& s) {
  const F3 e = sub(b, a), w = sub(o, a);
  const float ee = dot3(e, e), de = dot3(d, e), den = ee - de * de;
  s = den > 1e-14f ? fminf(fmaxf((dot3(e, w) - de * dot3(d, w)) / den, 0.f), 1.f) : 0.f;
  const F3 p = add(a, scale(s, e));
  const float t = fmaxf(dot3(sub(p, o), d), 0.f);
  const float miss = norm(sub(add(o, scale(t, d)), p));
  if (miss > rad) return -1.f;
  return t - sqrtf(rad * rad - miss * miss);
}

__device__ __forceinline__ unsigned long long key64(float d, uint32_t n) {
  return ((unsigned long long)__float_as_uint(d) << 32) | n;  // d >= 0: its bits order like its value
}

// the nearest hit among node n's primitives, as a key (kNoKey: none)
__device__ unsigned long long surface_key(const Ctx& c, uint32_t n) {
  const F3 o = f3(c.q.eye), d = f3(c.q.ray);
  const Cell L = cell_of(c, n);
  unsigned long long best = kNoKey;
  const bool intact = cell_intact(c, n);
  if (intact) {
    const uint32_t v00 = n, v10 = n + 1, v01 = n + L.nx, v11 = n + L.nx + 1;
    const F3 p00 = at(c, v00), p10 = at(c, v10), p01 = at(c, v01), p11 = at(c, v11);
    // the cell as triangles (00, 10, 11) and (00, 11, 01); the hit picks its heaviest corner
    float u, v;
    const float t1 = ray_triangle(o, d, p00, p10, p11, u, v);
    if (t1 > 0.f) {
      const float w0 = 1.f - u - v;
      const uint32_t m = (w0 >= u && w0 >= v) ? v00 : (u >= v ? v10 : v11);
      best = min(best, key64(t1, m));
    }
    const float t2 = ray_triangle(o, d, p00, p11, p01, u, v);
    if (t2 > 0.f) {
      const float w0 = 1.f - u - v;
      const uint32_t m = (w0 >= u && w0 >= v) ? v00 : (u >= v ? v11 : v01);
      best = min(best, key64(t2, m));
    }
  }
  if (!cloth(c, n)) return best;
  for (int fam = 0; fam < 2; fam++) {
    const uint32_t m = fam == 0 ? n + 1 : n + L.nx;
    if (!span_alive(c, n, fam) || !cloth(c, m)) continue;
    // a strand bordering an intact cell is picked through that cell
    const bool beside = fam == 0 ? (L.j > 0 && cell_intact(c, n - L.nx)) : (L.i > 0 && cell_intact(c, n - 1));
    if (intact || beside) continue;
    float sp;
    const float t = ray_strand(o, d, at(c, n), at(c, m), c.q.strand_radius, sp);
    if (t > 0.f) best = min(best, key64(t, sp > 0.5f ? m : n));
  }
  return best;
}

__global__ __launch_bounds__(256) void k_pick_surface(Ctx c) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)c.N) return;
  const unsigned long long k = surface_key(c, n);
  if (k != kNoKey) atomicMin(reinterpret_cast<unsigned long long*>(c.work), k);
}

// screen position (pixels x, y), NDC depth and clip w of a point
__device__ void to_screen(const Ctx& c, F3 p, float s[4]) {
  const float* M = c.q.view_proj;  // column-major
  float h[4];
  for (int r = 0; r < 4; r++) h[r] = M[r] * p.x + M[4 + r] * p.y + M[8 + r] * p.z + M[12 + r];
  s[0] = (0.5f + 0.5f * h[0] / h[3]) * c.q.width;
  s[1] = (0.5f - 0.5f * h[1] / h[3]) * c.q.height;
  s[2] = h[2] / h[3];
  s[3] = h[3];
}

__device__ __forceinline__ bool surface_hit(const Ctx& c) {
  return *reinterpret_cast<const unsigned long long*>(c.work) != kNoKey;
}

// crossings with material inside the cursor disc, compacted per wave
__global__ __launch_bounds__(256) void k_pick_disc(Ctx c) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (surface_hit(c)) return;
  bool in = false;
  if (n < (uint32_t)c.N && has_material(c, n)) {
    float s[4];
    to_screen(c, at(c, n), s);
    const float px = s[0] - c.q.pointer[0], py = s[1] - c.q.pointer[1];
    in = s[3] > 0.f && s[2] > 0.f && s[2] < 1.f && px * px + py * py <= c.q.radius_px * c.q.radius_px;
  }
  const unsigned long long mask = __ballot(in);
  if (mask == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(&c.work[4], (uint32_t)__popcll(mask));
  base = __shfl(base, leader, 64);
  if (in) {
    const uint32_t slot = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    if (slot < (uint32_t)kCand) c.work[5 + slot] = n;
  }
}

// every intact cell whose screen box (grown by the disc radius) holds the pointer marks the
// candidates it stands in front of
__global__ __launch_bounds__(256) void k_pick_hide(Ctx c) {
  const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= (uint32_t)c.N || surface_hit(c)) return;
  const uint32_t count = min(c.work[4], (uint32_t)kCand);
  if (count == 0 || !cell_intact(c, f)) return;
  const Cell L = cell_of(c, f);
  const uint32_t v[4] = {f, f + 1, f + L.nx, f + L.nx + 1};
  F3 p[4];
  float lo[2] = {3.4e38f, 3.4e38f}, hi[2] = {-3.4e38f, -3.4e38f};
  for (int q = 0; q < 4; q++) {
    p[q] = at(c, v[q]);
    float s[4];
    to_screen(c, p[q], s);
    if (s[3] <= 0.f) return;
    lo[0] = fminf(lo[0], s[0]);
    lo[1] = fminf(lo[1], s[1]);
    hi[0] = fmaxf(hi[0], s[0]);
    hi[1] = fmaxf(hi[1], s[1]);
  }
  const float R = c.q.radius_px, px = c.q.pointer[0], py = c.q.pointer[1];
  if (px < lo[0] - R || px > hi[0] + R || py < lo[1] - R || py > hi[1] + R) return;
  const F3 eye = f3(c.q.eye);
  for (uint32_t k = 0; k < count; k++) {
    const uint32_t m = c.work[5 + k];
    if (m == v[0] || m == v[1] || m == v[2] || m == v[3]) continue;
    const F3 to = sub(at(c, m), eye);
    const float l = norm(to);
    const F3 d = scale(1.f / l, to);
    float u, w;
    const float t1 = ray_triangle(eye, d, p[0], p[1], p[3], u, w);
    const float t2 = ray_triangle(eye, d, p[0], p[3], p[2], u, w);
    if ((t1 > 0.f && t1 < l - kBehind) || (t2 > 0.f && t2 < l - kBehind)) atomicOr(&c.work[5 + kCand + k], 1u);
  }
}

// the visible candidate nearest the pointer on screen
__global__ __launch_bounds__(128) void k_pick_visible(Ctx c) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (surface_hit(c) || k >= min(c.work[4], (uint32_t)kCand) || c.work[5 + kCand + k]) return;
  const uint32_t n = c.work[5 + k];
  float s[4];
  to_screen(c, at(c, n), s);
  const float px = s[0] - c.q.pointer[0], py = s[1] - c.q.pointer[1];
  atomicMin(reinterpret_cast<unsigned long long*>(c.work + 2), key64(sqrtf(px * px + py * py), n));
}

// the crossing with material nearest to a point within a radius (staged grabs)
__global__ __launch_bounds__(256) void k_pick_nearest(Ctx c) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)c.N || !has_material(c, n)) return;
  const float d = norm(sub(at(c, n), f3(c.q.eye)));
  if (d <= c.q.strand_radius) atomicMin(reinterpret_cast<unsigned long long*>(c.work + 2), key64(d, n));
}

// the result (point, node, kind 1 surface / 2 proximity / 0 none) and the connectivity window around
// the node: per lattice offset in [-R, R]^2, bit 0 / 1 its own u / v span survives, bit 2 not cloth,
// bit 3 outside the lattice
__global__ __launch_bounds__(256) void k_pick_result(Ctx c) {
  const unsigned long long surf = *reinterpret_cast<const unsigned long long*>(c.work);
  const unsigned long long near = *reinterpret_cast<const unsigned long long*>(c.work + 2);
  uint32_t node = kNoNode, kind = 0;
  F3 p = {0.f, 0.f, 0.f};
  if (surf != kNoKey) {
    node = (uint32_t)(surf & 0xFFFFFFFFull);
    p = add(f3(c.q.eye), scale(__uint_as_float((uint32_t)(surf >> 32)), f3(c.q.ray)));
    kind = 1;
  } else if (near != kNoKey) {
    node = (uint32_t)(near & 0xFFFFFFFFull);
    p = at(c, node);
    kind = 2;
  }
  const int t = threadIdx.x;
  if (t == 0) {
    c.out[0] = __float_as_uint(p.x);
    c.out[1] = __float_as_uint(p.y);
    c.out[2] = __float_as_uint(p.z);
    c.out[3] = node;
    c.out[4] = kind;
  }
  const int R = c.q.window_radius, side = 2 * R + 1;
  if (t >= side * side) return;
  uint32_t b = 8u;
  if (kind != 0) {
    const Cell L = cell_of(c, node);
    const int i = L.i + t % side - R, j = L.j + t / side - R;
    if (i >= 0 && j >= 0 && i < L.nx && j < L.ny) {
      const uint32_t m = (uint32_t)(L.off + j * L.nx + i);
      b = (span_alive(c, m, 0) ? 1u : 0u) | (span_alive(c, m, 1) ? 2u : 0u) | (cloth(c, m) ? 0u : 4u);
    }
  }
  c.out[8 + t] = b;
}

Ctx context(Solver& s) {
  Ctx c;
  c.N = s.N;
  c.Np = s.Np;
  c.geo = s.geo;
  c.x = s.x;
  c.dx = s.dx;
  c.state = s.nstate;
  c.flags = s.flags;
  c.slots = s.grip_slots;
  c.entries = s.grip_entries;
  c.handle = s.grip_handle;
  c.ext = s.ext;
  c.accumulate = s.contact_mode != kContactNone ? 1 : 0;
  c.q = s.pick_args;
  c.work = s.pick;
  c.out = s.pick_out;
  return c;
}

}  // namespace

void interact_alloc(Solver& s) {
  if (s.grip_slots) return;
  BD_HIP(hipMalloc(&s.grip_slots, kGrips * sizeof(GripSlot)));
  BD_HIP(hipMalloc(&s.grip_entries, (size_t)kGrips * kSlotEntries * sizeof(GripEntry)));
  BD_HIP(hipMalloc(&s.grip_handle, 12 * kGrips * sizeof(float)));
  BD_HIP(hipMalloc(&s.pick, kPickWords * sizeof(uint32_t)));
  BD_HIP(hipMalloc(&s.pick_out, (8 + 256) * sizeof(uint32_t)));
  BD_HIP(hipMemsetAsync(s.grip_slots, 0, kGrips * sizeof(GripSlot), s.stream));
  BD_HIP(hipMemsetAsync(s.grip_entries, 0, (size_t)kGrips * kSlotEntries * sizeof(GripEntry), s.stream));
  BD_HIP(hipMemsetAsync(s.grip_handle, 0, 12 * kGrips * sizeof(float), s.stream));
}

void launch_grip_init(Solver& s, int g) {
  const int count = s.grip_slots_h[g].count;
  k_grip_offsets<<<grid_for(count > 0 ? count : 1, 64), 64, 0, s.stream>>>(context(s), g);
}

void launch_grip_lead(Solver& s) {
  if (!s.grips_active) return;
  ProfScope ps(s, PC_PREDICT_COMMIT, 0.0);
  k_grip_wave<<<kGrips, 64, 0, s.stream>>>(context(s));
}

void launch_grip_terms(Solver& s) {
  if (!s.grips_active) return;
  ProfScope ps(s, PC_CONTACT_TERMS, 0.0);
  k_grip_springs<<<kGrips * kSlotEntries / 64, 64, 0, s.stream>>>(context(s));
}

void launch_pick(Solver& s, bool ray) {
  const Ctx c = context(s);
  // the keys start at "none", the candidate list empty
  std::vector<uint32_t> init(kPickWords, 0u);
  init[0] = init[1] = init[2] = init[3] = kNoNode;
  BD_HIP(hipMemcpyAsync(s.pick, init.data(), init.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s.stream));
  const int g = grid_for(s.N, 256);
  if (ray) {
    k_pick_surface<<<g, 256, 0, s.stream>>>(c);
    k_pick_disc<<<g, 256, 0, s.stream>>>(c);
    k_pick_hide<<<g, 256, 0, s.stream>>>(c);
    k_pick_visible<<<grid_for(kCand, 128), 128, 0, s.stream>>>(c);
  } else {
    k_pick_nearest<<<g, 256, 0, s.stream>>>(c);
  }
  k_pick_result<<<1, 256, 0, s.stream>>>(c);
}

}  // namespace bd

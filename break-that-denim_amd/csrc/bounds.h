// Contact bounds on a node's step and the floor law, shared by the contact, assembly, line-search
// and update kernels. Restates step.wgsl:61-82 (bound, feasible, floorBound) and
// common.wgsl:247-265 (floorLowest, floorWeightDepth) with FLOOR_CONTACT (src/simulation/floor.ts).
#pragma once
#include "bd_internal.h"

namespace bd {

constexpr float kUnbounded = 1e9f;
// FLOOR_CONTACT: contact distance beyond half the thickness, penalty k (N/m), friction, and the
// speed below which friction is viscous (m/s)
constexpr float kFloorBulge = 0.3e-3f;
constexpr float kFloorK = 1300.f;
constexpr float kFloorMu = 0.35f;
constexpr float kFloorSpeed = 0.002f;
// stiffness of an active bound relative to the node's own diagonal (assemble.wgsl:139-142)
constexpr float kActiveBound = 100.f;

// a half-space bound: the step may move against m.xyz by at most m.w
__device__ __forceinline__ void bound_step(float d[3], float4 m) {
#pragma clang fp contract(off)
  float approach = -((d[0] * m.x + d[1] * m.y) + d[2] * m.z);
  if (approach > m.w) {
    float e = approach - m.w;
    d[0] = d[0] + e * m.x;
    d[1] = d[1] + e * m.y;
    d[2] = d[2] + e * m.z;
  }
}

// own, reciprocal, own: alternating projections onto the two half-spaces
__device__ __forceinline__ void feasible_step(float d[3], float4 own, float4 rec) {
  bound_step(d, own);
  bound_step(d, rec);
  bound_step(d, own);
}

// lowest height an iterate may reach from zLin in one update (common.wgsl floorLowest)
__device__ __forceinline__ float floor_lowest(float zLin, float fz, float delta, float half, float sWeight) {
#pragma clang fp contract(off)
  float s = fmaxf(delta - (zLin - fz), 0.f);
  return fminf(fmaxf((fz + delta) - fmaxf(sWeight, 2.f * s), fz + half), zLin);
}

// depth into the contact distance at which the penalty carries a node of mass m (floorWeightDepth)
__device__ __forceinline__ float floor_weight_depth(float m, float g, float delta, float k) {
#pragma clang fp contract(off)
  return sqrtf(((m * g) * delta) / k);
}

// the step d's z bounded against the floor from the iterate step linz (step.wgsl floorBound)
__device__ __forceinline__ float floor_bound_z(float z0, float linz, float dz, float half, float mass, float gnorm, float fz) {
#pragma clang fp contract(off)
  float delta = half + kFloorBulge;
  float sW = floor_weight_depth(mass, gnorm, delta, kFloorK);
  float lowest = floor_lowest(z0 + linz, fz, delta, half, sW);
  return fmaxf(dz, lowest - z0);
}

}  // namespace bd

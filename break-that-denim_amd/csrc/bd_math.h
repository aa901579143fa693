// Device math shared by the kernels: symmetric 3x3 blocks, the constitutive laws, constants.
//
// The laws restate src/simulation/axial.ts:16-85 and src/simulation/laws.ts:18-159 (mirrored by
// the reference's common.wgsl); they are templated on the arithmetic type so the solve can run
// them in FP32 and the history updates and line-search energies in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace bd {

constexpr double kBendS0 = 0.85;
constexpr double kShearJam = 0.16;
constexpr double kShearRamp = 0.08;
constexpr double kShearLock = 40.0;
// sqrt(1 - (1 - (1 - SHEAR_JAM)^2)^2) (laws.ts:135)
constexpr double kShearCos0 = 0.9556822903036343;
constexpr double kPivotShift = 2e-6;
constexpr double kLineNoise = 4e-6;
constexpr double kLinkMinIntact = 0.02;
constexpr double kAreaMin = 0.02;

constexpr uint16_t SEG_EXISTS = 1;
constexpr uint16_t SEG_BROKEN = 2;
constexpr uint16_t SEG_SLIPPED = 8;
constexpr uint8_t DAMAGE_CUT = 255;
constexpr uint8_t DAMAGE_PULLED = 254;
constexpr uint8_t DAMAGE_SEPARATED = 254;

struct DevYarn {
  double lin, ea_low, ea_high, crimp, ea_comp, visc, bend_rig, bend_visc, strength, soft;
  int B, pad;
};

struct DevLayer {
  int nx, ny, off, bu, bv, foff, fk, fm;
  double wpu, wpv;   // individual-yarn pitches
  double pu, pv;     // mechanical pitches
  DevYarn y[2];
  double shear_k, mu, cn0, wrap, air_drag, thickness;
};

// ---------------------------------------------------------------- symmetric 3x3 blocks
// storage order [xx, xy, xz, yy, yz, zz]
struct Sym {
  float a[6];
};

__device__ __forceinline__ void sym_zero(Sym& s) {
#pragma unroll
  for (int q = 0; q < 6; q++) s.a[q] = 0.f;
}

__device__ __forceinline__ void sym_add_diag(Sym& s, float k) {
  s.a[0] += k;
  s.a[3] += k;
  s.a[5] += k;
}

// s += c * [(kpar - kperp) e e^T + kperp I]
__device__ __forceinline__ void sym_add_axial(Sym& s, float c, float kpar, float kperp, const float e[3]) {
  float d = c * (kpar - kperp);
  float kp = c * kperp;
  s.a[0] += d * e[0] * e[0] + kp;
  s.a[1] += d * e[0] * e[1];
  s.a[2] += d * e[0] * e[2];
  s.a[3] += d * e[1] * e[1] + kp;
  s.a[4] += d * e[1] * e[2];
  s.a[5] += d * e[2] * e[2] + kp;
}

__device__ __forceinline__ void sym_mul(const float* s, const float v[3], float o[3]) {
  o[0] = s[0] * v[0] + s[1] * v[1] + s[2] * v[2];
  o[1] = s[1] * v[0] + s[3] * v[1] + s[4] * v[2];
  o[2] = s[2] * v[0] + s[4] * v[1] + s[5] * v[2];
}

__device__ __forceinline__ float sym_frob(const float* s) {
  return sqrtf(s[0] * s[0] + s[3] * s[3] + s[5] * s[5] + 2.f * (s[1] * s[1] + s[2] * s[2] + s[4] * s[4]));
}

// ---------------------------------------------------------------- laws
template <class T>
__device__ __forceinline__ T axial_tension(const DevYarn& p, T eps) {
  if (eps <= T(0)) return T(p.ea_comp) * eps;
  T ec = T(p.crimp), dk = T(p.ea_high) - T(p.ea_low);
  if (eps <= ec) return T(p.ea_low) * eps + (dk * eps * eps) / (T(2) * ec);
  return T(p.ea_low) * ec + (dk * ec) / T(2) + T(p.ea_high) * (eps - ec);
}

template <class T>
__device__ __forceinline__ T axial_stiffness(const DevYarn& p, T eps) {
  if (eps <= T(0)) return T(p.ea_comp);
  T ec = T(p.crimp);
  if (eps <= ec) return T(p.ea_low) + ((T(p.ea_high) - T(p.ea_low)) * eps) / ec;
  return T(p.ea_high);
}

template <class T>
__device__ __forceinline__ T axial_energy_density(const DevYarn& p, T eps) {
  if (eps <= T(0)) return T(0.5) * T(p.ea_comp) * eps * eps;
  T ec = T(p.crimp), dk = T(p.ea_high) - T(p.ea_low);
  if (eps <= ec) return T(0.5) * T(p.ea_low) * eps * eps + (dk * eps * eps * eps) / (T(6) * ec);
  T tc = T(p.ea_low) * ec + (dk * ec) / T(2);
  T ece = T(0.5) * T(p.ea_low) * ec * ec + (dk * ec * ec) / T(6);
  T d = eps - ec;
  return ece + tc * d + T(0.5) * T(p.ea_high) * d * d;
}

template <class T>
__device__ __forceinline__ T strain_at_tension(const DevYarn& p, T t) {
  T ec = T(p.crimp), dk = T(p.ea_high) - T(p.ea_low);
  T tc = T(p.ea_low) * ec + (dk * ec) / T(2);
  if (t >= tc) return ec + (t - tc) / T(p.ea_high);
  T a = dk / (T(2) * ec);
  return (T(2) * t) / (T(p.ea_low) + sqrt(T(p.ea_low) * T(p.ea_low) + T(4) * a * t));
}

// [Psi, Psi', Psi''] of the fold-stiffening bending (laws.ts:18-30)
template <class T>
__device__ __forceinline__ void bend_psi(T q, T a, T& p0, T& p1, T& p2) {
  T s = q / a;
  if (s < T(kBendS0)) {
    T d = T(1) - s;
    p0 = q / d;
    p1 = T(1) / (d * d);
    p2 = T(2) / (a * d * d * d);
    return;
  }
  T d0 = T(1) - T(kBendS0), q0 = T(kBendS0) * a;
  T c1 = T(1) / (d0 * d0), c2 = T(2) / (a * d0 * d0 * d0), t = q - q0;
  p0 = q0 / d0 + c1 * t + T(0.5) * c2 * t * t;
  p1 = c1 + c2 * t;
  p2 = c2;
}

__device__ __forceinline__ double bend_psi_delta(double q0, double q1, double dq, double a) {
  double s0 = q0 / a, s1 = q1 / a;
  if (s0 < kBendS0 && s1 < kBendS0) return dq / ((1 - s0) * (1 - s1));
  double qb = kBendS0 * a, d0 = 1 - kBendS0;
  double p1 = 1 / (d0 * d0), p2 = 2 / (a * d0 * d0 * d0);
  if (s0 >= kBendS0 && s1 >= kBendS0) return dq * (p1 + 0.5 * p2 * (q0 - qb + (q1 - qb)));
  double lo = fmin(q0, q1);
  double t = fmax(q0, q1) - qb;
  double up = (qb - lo) / ((1 - lo / a) * d0) + t * (p1 + 0.5 * p2 * t);
  return q1 >= q0 ? up : -up;
}

// bendGeometry (laws.ts:52-56): [wa, wc, kPerB, fold]
template <class T>
__device__ __forceinline__ void bend_geometry(T La, T Lc, T& wa, T& wc, T& kperb, T& fold) {
  T s = La + Lc;
  T lh = (T(2) * La * Lc) / s;
  wa = (T(2) * Lc) / s;
  wc = (T(2) * La) / s;
  kperb = T(1) / (lh * lh * T(0.5) * s);
  fold = T(4) * lh * lh;
}

template <class T>
__device__ __forceinline__ T bend_scale(T ia, T ic, uint32_t ma, uint32_t mc, int B) {
  uint32_t full = B >= 32 ? 0xffffffffu : ((1u << B) - 1u);
  int cont = __popc(~(ma | mc) & full);
  if (cont == 0) return T(0);
  int ua = __popc(~ma & full), uc = __popc(~mc & full);
  T ra = ua > 0 ? (ia * T(B)) / T(ua) : T(0);
  T rc = uc > 0 ? (ic * T(B)) / T(uc) : T(0);
  return (T(cont) / T(B)) * (ra < rc ? ra : rc);
}

template <class T>
__device__ __forceinline__ T shear_jam(T area) {
  T ar = area < T(1) ? area : T(1);
  ar = ar > T(0.02) ? ar : T(0.02);
  T c = T(kShearCos0) * sqrt(ar);
  return T(1) - sqrt(T(1) - sqrt(T(1) - c * c));
}

template <class T>
__device__ __forceinline__ T cell_area(T cu, T tu, T cv, T rv) {
  T a = T(0.5) * (cu + tu), b = T(0.5) * (cv + rv);
  return a > b ? a : b;
}

// shearLaw (laws.ts:96-115): force, tangent, energy
template <class T>
__device__ __forceinline__ void shear_law(T ks, T Ld, T l, T jam, T& f, T& k, T& e) {
  T d = l - Ld;
  f = ks * d;
  k = ks;
  e = T(0.5) * ks * d * d;
  T c = -d / Ld - jam;
  if (c > T(0)) {
    T extra = ks * T(kShearLock - 1);
    T R = T(kShearRamp);
    if (c < R) {
      f -= (extra * Ld * c * c) / (T(2) * R);
      k += (extra * c) / R;
      e += (extra * Ld * Ld * c * c * c) / (T(6) * R);
    } else {
      f -= extra * Ld * (c - T(0.5) * R);
      k += extra;
      e += extra * Ld * Ld * ((R * R) / T(6) + T(0.5) * (c - R) * (c - R) + T(0.5) * R * (c - R));
    }
  }
}

// ---------------------------------------------------------------- lattice geometry
struct Geo {
  int nl;
  int nx[8], ny[8], off[8];
};

__device__ __forceinline__ int geo_layer(const Geo& g, int n) {
  int l = 0;
  for (int q = 1; q < g.nl; q++)
    if (n >= g.off[q]) l = q;
  return l;
}

// rotationCurve (common.wgsl:277-283): a rotation by th (angle |th| about th / |th|) moves the arm
// r by (R - I) r; its tangent is th x r, and this is what the rotation adds to it:
// (sin phi - phi) k x r + (1 - cos phi) k x (k x r).
__device__ __forceinline__ void rotation_curve(const double th[3], const double r[3], double o[3]) {
  const double phi = sqrt(th[0] * th[0] + th[1] * th[1] + th[2] * th[2]);
  if (phi < 1e-9) {
    o[0] = o[1] = o[2] = 0;
    return;
  }
  const double k[3] = {th[0] / phi, th[1] / phi, th[2] / phi};
  const double kr[3] = {k[1] * r[2] - k[2] * r[1], k[2] * r[0] - k[0] * r[2], k[0] * r[1] - k[1] * r[0]};
  const double kkr[3] = {k[1] * kr[2] - k[2] * kr[1], k[2] * kr[0] - k[0] * kr[2], k[0] * kr[1] - k[1] * kr[0]};
  const double a = sin(phi) - phi, c = 1.0 - cos(phi);
  for (int q = 0; q < 3; q++) o[q] = a * kr[q] + c * kkr[q];
}

// XCD-aware block index: consecutive logical tiles land on the same XCD (8 XCDs, blocks are
// dispatched round-robin across them), so neighbouring rows share that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
  const int X = 8;
  int per = nblocks / X;
  if (per == 0 || b >= per * X) return b;
  return (b % X) * per + (b / X);
}

}  // namespace bd

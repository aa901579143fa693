// History updates after commit: the per-yarn damage field (cohesive damage with load sharing,
// shear-lag pull-out, span coupling and rupture) and friction-limited crossing slip.
//
// Restates updateDamage / updateSlip (src/simulation/reference/solver.ts:659-740) and
// updateField / pullOut / coupleSpans / partSpanYarns (src/simulation/yarnField.ts:125-499).
// Everything here is FP64 and this file is compiled with -ffp-contract=off, with every expression in
// the oracle's operation order, so damage codes, rupture and pull-out events are bit-exact against
// the FP64 CPU oracle on the same committed positions. Each pass has a single writer per output:
// a (span, yarn) owns its crossings, a span its two tail slots, a slip parity pass its spans.
#include <stdexcept>

#include "bd_internal.h"

#pragma clang fp contract(off)

namespace bd {

namespace {

struct HistArgs {
  int N, Np, C, nl;
  Geo geo;
  const DevLayer* layers;
  const float* x;
  const uint8_t* pinned;
  const double* cn0;
  double *rest, *intact, *fail_site, *tails, *strain;
  const double* softening;
  uint32_t* cut;
  uint16_t* flags;
  uint8_t *damage, *sep, *mark;
  const uint8_t* area;
  const float* strength;
  int* counters;
};

__device__ __forceinline__ bool h_active(const HistArgs& A, int s) { return s >= 0 && (A.flags[s] & (SEG_EXISTS | SEG_BROKEN)) == SEG_EXISTS; }

__device__ __forceinline__ double area_of(uint8_t c) { return fmax(kAreaMin, c / 255.0); }

__device__ __forceinline__ int bundle_of(const DevLayer& L, int f) { return f == 1 ? L.bu : L.bv; }
__device__ __forceinline__ int span_len(const DevLayer& L, int f) { return f == 1 ? L.bv : L.bu; }

__device__ __forceinline__ int span_crossing(const DevLayer& L, int i, int j, int f, int q, int p) {
  if (f == 1) return L.foff + (j * L.bv + L.bv / 2 + p) * L.fk + i * L.bu + q;
  return L.foff + (j * L.bv + q) * L.fk + i * L.bu + L.bu / 2 + p;
}

__device__ double seg_length(const HistArgs& A, int s, int b) {
  int a = s >> 1;
  const int Np = A.Np;
  double d0 = (double)A.x[b] - (double)A.x[a];
  double d1 = (double)A.x[Np + b] - (double)A.x[Np + a];
  double d2 = (double)A.x[2 * Np + b] - (double)A.x[2 * Np + a];
  return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

__device__ __forceinline__ void locate(const Geo& g, int n, int& l, int& i, int& j) {
  l = geo_layer(g, n);
  int r = n - g.off[l];
  j = r / g.nx[l];
  i = r - j * g.nx[l];
}

// other endpoint of span s, or -1 where the span cannot exist
__device__ __forceinline__ int seg_other(const HistArgs& A, int s) {
  int n = s >> 1, f = s & 1, l, i, j;
  locate(A.geo, n, l, i, j);
  if (f == 0) return i + 1 < A.geo.nx[l] ? n + 1 : -1;
  return j + 1 < A.geo.ny[l] ? n + A.geo.nx[l] : -1;
}

// strain snapshot (solver.ts:661)
__global__ __launch_bounds__(256) void k_strain(HistArgs A) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * A.N) return;
  int b = seg_other(A, s);
  // retired nodes are out of the scene (yarnfield.wgsl:50, 60)
  bool cloth = b >= 0 && !((A.pinned[s >> 1] | A.pinned[b]) & kPinRetired);
  A.strain[s] = (cloth && h_active(A, s)) ? seg_length(A, s, b) / A.rest[s] - 1 : __builtin_nan("");
}

__global__ __launch_bounds__(256) void k_sep(HistArgs A) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 2 * A.C) return;
  A.sep[q] = A.damage[q] >= DAMAGE_SEPARATED ? 1 : 0;
}

// interpolatedStrain (yarnField.ts:216-240)
__device__ double interp_strain(const DevLayer& L, int f, int k, int m, const double* strain) {
  const int bu = L.bu, bv = L.bv;
  double s = f == 1 ? (k - (bu - 1) / 2.0) / bu : (k - (bu - 0.5)) / bu;
  double t = f == 1 ? (m - (bv - 0.5)) / bv : (m - (bv - 1) / 2.0) / bv;
  double i0 = floor(s), j0 = floor(t);
  double fu = s - i0, fv = t - j0;
  int imax = f == 1 ? L.nx - 1 : L.nx - 2;
  int jmax = f == 1 ? L.ny - 2 : L.ny - 1;
  double num = 0, den = 0;
  for (int dj = 0; dj < 2; dj++)
    for (int di = 0; di < 2; di++) {
      int i = (int)i0 + di, j = (int)j0 + dj;
      if (i < 0 || j < 0 || i > imax || j > jmax) continue;
      double e = strain[2 * (L.off + j * L.nx + i) + f];
      if (!isfinite(e)) continue;
      double w = (di ? fu : 1 - fu) * (dj ? fv : 1 - fv);
      num += w * e;
      den += w;
    }
  return den > 1e-6 ? num / den : __builtin_nan("");
}

__device__ int share_reach(const DevLayer& L, int f) {
  double r = 1.2e-3 / (f == 1 ? L.wpv : L.wpu);
  int v = (int)floor(r + 0.5);
  return v > 1 ? v : 1;
}

// shareFactor (yarnField.ts:303-321)
__device__ double share_factor(const DevLayer& L, const uint8_t* sep, int f, int k, int m, int reach, double cap) {
  const int K = L.fk, M = L.fm;
  int along = f == 1 ? m : k, lenAlong = f == 1 ? M : K;
  int across = f == 1 ? k : m, lenAcross = f == 1 ? K : M;
  int runs = 0;
  for (int dir = -1; dir <= 1; dir += 2)
    for (int d = 1; d <= 6; d++) {
      int y = across + dir * d;
      if (y < 0 || y >= lenAcross) break;
      bool broken = false;
      int a0 = along - reach < 0 ? 0 : along - reach;
      int a1 = along + reach > lenAlong - 1 ? lenAlong - 1 : along + reach;
      for (int a = a0; a <= a1 && !broken; a++) {
        int at = L.foff + (f == 1 ? a * K + y : y * K + a);
        broken = sep[2 * at + f] == 1;
      }
      if (!broken) break;
      runs++;
    }
  return 1 + fmin((0.6 * runs) / 2, cap);
}

__device__ double yarn_damage(const DevYarn& p, double yarns, double area, double peak, double soft, double eps) {
  double ep = strain_at_tension<double>(p, (peak * yarns) / area);
  if (eps <= ep) return 0;
  if (eps >= ep + soft) return 1;
  double tenv = peak * (1 - (eps - ep) / soft);
  double tel = (area * axial_tension<double>(p, eps)) / yarns;
  double d = 1 - tenv / fmax(tel, 1e-12);
  return fmin(1.0, fmax(0.0, d));
}

__device__ __forceinline__ int damage_code(double d) {
  if (d >= 1) return DAMAGE_CUT;
  double c = floor(d * DAMAGE_CUT);
  return c < DAMAGE_PULLED - 1 ? (int)c : DAMAGE_PULLED - 1;
}

// updateField (yarnField.ts:338-410): one thread per (span of family f, yarn q)
__global__ __launch_bounds__(256) void k_update_field(HistArgs A, int f) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  int q = blockIdx.y;
  if (n >= A.N) return;
  int l, i, j;
  locate(A.geo, n, l, i, j);
  const DevLayer& L = A.layers[l];
  if ((f == 0 && i >= L.nx - 1) || (f == 1 && j >= L.ny - 1)) return;
  const int s = 2 * n + f;
  if (!h_active(A, s)) return;
  if ((A.pinned[n] | A.pinned[f == 0 ? n + 1 : n + L.nx]) & kPinRetired) return;  // yarnfield.wgsl:173
  const DevYarn& y = L.y[f];
  const int B = bundle_of(L, f), P = span_len(L, f), K = L.fk;
  if (q >= B) return;
  if ((A.cut[s] >> q) & 1) return;
  double soft = A.softening[s] / A.rest[s];
  int reach = share_reach(L, f);
  int mid = span_crossing(L, i, j, f, q, P >> 1) - L.foff;
  double eps = interp_strain(L, f, mid % K, mid / K, A.strain);
  if (!isfinite(eps)) return;
  double minPeak = __builtin_inf();
  for (int p = 0; p < P; p++) {
    int c = span_crossing(L, i, j, f, q, p);
    double pk = strain_at_tension<double>(y, ((double)A.strength[2 * c + f] * B) / area_of(A.area[2 * c + f]));
    if (pk < minPeak) minPeak = pk;
  }
  if (eps * (1 + 0.6 * 6) <= minPeak) return;
  int g = 1 - f;
  int kM = mid % K, mM = mid / K;
  double ec = interp_strain(L, g, kM, mM, A.strain);
  double tCross = isfinite(ec) ? (fmax(0.0, axial_tension<double>(L.y[g], ec)) * area_of(A.area[2 * (mid + L.foff) + g])) / bundle_of(L, g) : 0;
  double n0 = (double)A.cn0[s >> 1];
  double transfer = ((2 * reach + 1) * L.mu * (n0 / (L.bu * L.bv) + L.wrap * tCross)) / 1.0;
  int best = -1, soft0 = -1, softCode = 0;
  double bestScore = -__builtin_inf(), bestOmega = 1;
  for (int p = 0; p < P; p++) {
    int c = span_crossing(L, i, j, f, q, p);
    int local = c - L.foff;
    double omega = share_factor(L, A.sep, f, local % K, local / K, reach, transfer / A.strength[2 * c + f]);
    double peak = strain_at_tension<double>(y, ((double)A.strength[2 * c + f] * B) / area_of(A.area[2 * c + f]));
    double score = omega / peak;
    if (score > bestScore) {
      best = c;
      bestScore = score;
      bestOmega = omega;
    }
    if (A.damage[2 * c + f] > softCode) {
      soft0 = c;
      softCode = A.damage[2 * c + f];
    }
  }
  if (soft0 >= 0) {
    best = soft0;
    int local = best - L.foff;
    bestOmega = share_factor(L, A.sep, f, local % K, local / K, reach, transfer / A.strength[2 * best + f]);
  }
  int code = damage_code(yarn_damage(y, B, area_of(A.area[2 * best + f]), A.strength[2 * best + f], soft, eps * bestOmega));
  if (code > A.damage[2 * best + f]) A.damage[2 * best + f] = (uint8_t)code;
}

__device__ int break_distance(const DevLayer& L, const uint8_t* damage, int f, int k, int m) {
  int pos = f == 0 ? k : m, len = f == 0 ? L.fk : L.fm, stride = f == 0 ? 1 : L.fk;
  int c = L.foff + m * L.fk + k;
  int best = -1;
  for (int dir = -1; dir <= 1; dir += 2)
    for (int d = 1; d <= 12; d++) {
      int p = pos + dir * d;
      if (p < 0 || p >= len) break;
      int code = damage[2 * (c + dir * d * stride) + f];
      if (code == DAMAGE_CUT) {
        if (best < 0 || d < best) best = d;
        break;
      }
      if (code != DAMAGE_PULLED) break;
    }
  return best;
}

__device__ double recoil_length(const DevYarn& y, double B, double area, double peak, double eps, double mu, double n0, double wrap,
                                double tCross, double pitch) {
  if (!(eps > 0)) return 0;
  double share = area / B;
  double e = fmin(eps, strain_at_tension<double>(y, peak / share));
  double comp = share * (axial_tension<double>(y, e) * e - axial_energy_density<double>(y, e));
  return (comp * pitch) / (mu * (n0 + wrap * fmax(0.0, tCross)));
}

// pullOut marks (yarnField.ts:125-149): one thread per crossing
__global__ __launch_bounds__(256) void k_pull_mark(HistArgs A) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= A.C) return;
  int l = 0;
  for (int q = 1; q < A.nl; q++)
    if (c >= A.layers[q].foff) l = q;
  const DevLayer& L = A.layers[l];
  int local = c - L.foff;
  int m = local / L.fk, k = local - m * L.fk;
  for (int f = 0; f < 2; f++) {
    A.mark[2 * c + f] = 0;
    if (A.damage[2 * c + f] >= DAMAGE_SEPARATED) continue;
    int d = break_distance(L, A.damage, f, k, m);
    if (d < 0) continue;
    double eps = interp_strain(L, f, k, m, A.strain);
    if (!isfinite(eps)) continue;
    int g = 1 - f;
    double ec = interp_strain(L, g, k, m, A.strain);
    double tCross = isfinite(ec) ? (fmax(0.0, axial_tension<double>(L.y[g], ec)) * area_of(A.area[2 * c + g])) / bundle_of(L, g) : 0;
    int i = k / L.bu < L.nx - 1 ? k / L.bu : L.nx - 1;
    int j = m / L.bv < L.ny - 1 ? m / L.bv : L.ny - 1;
    int n = L.off + j * L.nx + i;
    double pitch = f == 1 ? L.wpv : L.wpu;
    double delta = recoil_length(L.y[f], bundle_of(L, f), area_of(A.area[2 * c + f]), A.strength[2 * c + f], eps, L.mu,
                                 (double)A.cn0[n] / (L.bu * L.bv), L.wrap, tCross, pitch);
    if (delta >= (d - 0.5) * pitch) A.mark[2 * c + f] = 1;
  }
}

__global__ __launch_bounds__(256) void k_pull_apply(HistArgs A) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 2 * A.C) return;
  if (A.mark[q]) A.damage[q] = DAMAGE_PULLED;
}

__device__ void part_span_yarns(const DevLayer& L, uint8_t* damage, int i, int j, int f, int slipEnd) {
  int P = span_len(L, f);
  int half = P - P / 2;
  for (int q = 0; q < bundle_of(L, f); q++) {
    bool ofA = slipEnd != 0;
    for (int p = 0; p < P; p++) {
      int e = 2 * span_crossing(L, i, j, f, q, p) + f;
      if (damage[e] >= DAMAGE_SEPARATED) {
        if (slipEnd < 0) ofA = false;
      } else if (ofA != (p < half)) {
        damage[e] = DAMAGE_PULLED;
      }
    }
  }
}

// coupleSpans (yarnField.ts:443-499): one thread per span
__global__ __launch_bounds__(256) void k_couple(HistArgs A) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 2 * A.N) return;
  int n = s >> 1, f = s & 1, l, i, j;
  locate(A.geo, n, l, i, j);
  const DevLayer& L = A.layers[l];
  if ((f == 0 && i >= L.nx - 1) || (f == 1 && j >= L.ny - 1)) return;
  if (A.flags[s] & SEG_SLIPPED) {
    part_span_yarns(L, A.damage, i, j, f, A.fail_site[s] < 0.5 ? 0 : 1);
    A.flags[s] &= (uint16_t)~SEG_SLIPPED;
    return;
  }
  if (!h_active(A, s)) return;
  if ((A.pinned[n] | A.pinned[f == 0 ? n + 1 : n + L.nx]) & kPinRetired) return;  // yarnfield.wgsl:397
  int B = bundle_of(L, f), P = span_len(L, f);
  double sum = 0, site = 0;
  uint32_t mask = 0;
  for (int q = 0; q < B; q++) {
    int dmax = 0, first = -1;
    double compliance = 0;
    for (int p = 0; p < P; p++) {
      int c = span_crossing(L, i, j, f, q, p);
      int d = A.damage[2 * c + f];
      if (d > dmax) dmax = d;
      if (d >= DAMAGE_SEPARATED && first < 0) first = p;
      compliance += 1 / area_of(A.area[2 * c + f]);
    }
    sum += (dmax >= DAMAGE_SEPARATED ? 0 : 1 - (double)dmax / DAMAGE_CUT) * (P / compliance);
    if (dmax >= DAMAGE_SEPARATED) {
      mask |= 1u << q;
      site += (first + 0.5) / P;
    }
  }
  A.intact[s] = sum / B;
  A.cut[s] = mask;
  uint32_t full = B == 32 ? 0xffffffffu : ((1u << B) - 1u);
  if (mask == full) {
    double phi = site / B;
    double rest = A.rest[s];
    int b = f == 0 ? n + 1 : n + L.nx;
    A.flags[s] |= SEG_BROKEN;
    A.fail_site[s] = phi;
    A.tails[4 * n + 2 * f + 1] += rest * phi;
    A.tails[4 * b + 2 * f] += rest * (1 - phi);
    atomicAdd(&A.counters[0], 1);
    part_span_yarns(L, A.damage, i, j, f, -1);
  }
}

// segTension (solver.ts:647-652)
__device__ void seg_tension(const HistArgs& A, const DevLayer& L, int s, int b, double& T, double& dTdL) {
  double Lr = A.rest[s], l = seg_length(A, s, b);
  double eps = l / Lr - 1;
  const DevYarn& y = L.y[s & 1];
  T = A.intact[s] * axial_tension<double>(y, eps);
  double K = A.intact[s] * axial_stiffness<double>(y, eps);
  dTdL = fmax(K, 0.0) * (l / (Lr * Lr));
}

// slipNode (solver.ts:685-740) for the nodes of one parity along family f
__global__ __launch_bounds__(256) void k_slip(HistArgs A, int f, int parity) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= A.N) return;
  int l, i, j;
  locate(A.geo, n, l, i, j);
  const DevLayer& L = A.layers[l];
  int along = f == 0 ? i : j;
  if ((along & 1) != parity) return;
  if (A.pinned[n]) return;
  const int nx = L.nx;
  int prev = f == 0 ? (i > 0 ? n - 1 : -1) : (j > 0 ? n - nx : -1);
  int next = f == 0 ? (i + 1 < nx ? n + 1 : -1) : (j + 1 < L.ny ? n + nx : -1);
  if (prev < 0 || next < 0) return;
  int sL = 2 * prev + f, sR = 2 * n + f;
  bool actL = h_active(A, sL), actR = h_active(A, sR);
  if (!actL && !actR) return;
  int tailL = 4 * n + 2 * f, tailR = 4 * n + 2 * f + 1;
  double TL = 0, kL = 0, TR = 0, kR = 0;
  if (actL) seg_tension(A, L, sL, n, TL, kL);
  if (actR) seg_tension(A, L, sR, next, TR, kR);
  double tc = 0;
  int cnt = 0;
  int g = 1 - f;
  int cp = g == 0 ? (i > 0 ? n - 1 : -1) : (j > 0 ? n - nx : -1);
  int cn = g == 0 ? (i + 1 < nx ? n + 1 : -1) : (j + 1 < L.ny ? n + nx : -1);
  if (cp >= 0 && h_active(A, 2 * cp + g)) {
    double T, k;
    seg_tension(A, L, 2 * cp + g, n, T, k);
    tc += fmax(0.0, T);
    cnt++;
  }
  if (cn >= 0 && h_active(A, 2 * n + g)) {
    double T, k;
    seg_tension(A, L, 2 * n + g, cn, T, k);
    tc += fmax(0.0, T);
    cnt++;
  }
  if (cnt == 0) return;
  tc /= cnt;
  double Fc = L.mu * ((double)A.cn0[n] + L.wrap * L.y[f].B * tc);
  double dT = fmax(TR, 0.0) - fmax(TL, 0.0);
  if (fabs(dT) <= Fc) return;
  bool highR = dT > 0;
  double kHigh = highR ? kR : kL;
  bool lowActive = highR ? actL : actR;
  double kLow = lowActive ? (highR ? kL : kR) : 0;
  double dl = (fabs(dT) - Fc) / fmax(kHigh + kLow, 1e-9);
  double pitch = f == 0 ? L.pu : L.pv;
  int sHigh = highR ? sR : sL, sLow = highR ? sL : sR;
  if (lowActive) {
    dl = fmin(dl, fmax(0.0, A.rest[sLow] - 0.55 * pitch));
    A.rest[sLow] -= dl;
    A.rest[sHigh] += dl;
  } else {
    int tIdx = highR ? tailL : tailR;
    dl = fmin(dl, A.tails[tIdx]);
    A.tails[tIdx] -= dl;
    A.rest[sHigh] += dl;
    if (A.tails[tIdx] <= 1e-9) {
      A.tails[tIdx] = 0;
      int far = highR ? next : prev;
      int farTail = highR ? 4 * far + 2 * f : 4 * far + 2 * f + 1;
      A.tails[farTail] += A.rest[sHigh];
      A.flags[sHigh] |= SEG_BROKEN | SEG_SLIPPED;
      A.fail_site[sHigh] = highR ? 0 : 1;
      atomicAdd(&A.counters[1], 1);
    }
  }
}

HistArgs hist_args(Solver& s) {
  return HistArgs{s.N, s.Np, s.C, s.nl, s.geo, s.layers, s.x, s.pinned, s.cn0, s.rest, s.intact, s.fail_site, s.tails, s.strain,
                  s.softening, s.cut_mask, s.flags, s.damage, s.sep, s.pullmark, s.area, s.strength, s.counters};
}

}  // namespace

// updateDamage (solver.ts:659-666): strain snapshot, field update, pull-out, span coupling
void launch_damage(Solver& s) {
  HistArgs A = hist_args(s);
  ProfScope ps(s, PC_DAMAGE, s.N * 100.0 + s.C * 24.0);
  k_strain<<<grid_for(2 * s.N, 256), 256, 0, s.stream>>>(A);
  if (s.C > 0) {
    k_sep<<<grid_for(2 * s.C, 256), 256, 0, s.stream>>>(A);
    int bmax = 1;
    for (int l = 0; l < s.nl; l++) bmax = std::max(bmax, std::max(s.layers_h[l].bu, s.layers_h[l].bv));
    for (int f = 0; f < 2; f++) k_update_field<<<dim3(grid_for(s.N, 256), bmax), 256, 0, s.stream>>>(A, f);
    k_pull_mark<<<grid_for(s.C, 256), 256, 0, s.stream>>>(A);
    k_pull_apply<<<grid_for(2 * s.C, 256), 256, 0, s.stream>>>(A);
  }
  k_couple<<<grid_for(2 * s.N, 256), 256, 0, s.stream>>>(A);
}

// updateSlip (solver.ts:669-683): 2 families x 2 parities
void launch_slip(Solver& s) {
  HistArgs A = hist_args(s);
  ProfScope ps(s, PC_SLIP, s.N * 4 * 60.0);
  for (int f = 0; f < 2; f++)
    for (int par = 0; par < 2; par++) k_slip<<<grid_for(s.N, 256), 256, 0, s.stream>>>(A, f, par);
}

}  // namespace bd

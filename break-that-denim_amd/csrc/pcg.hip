// PCG with the symmetric V-cycle preconditioner, and the line-search energy changes.
//
// pcg restates newtonMultigrid.ts:537-576: r = b - A d (d = 0), z = V(r), p = z, then per
// iteration Ap = A p, alpha = r.z / p.Ap, d += alpha p, r -= alpha Ap, z = V(r), beta, p = z + beta p;
// the last fixed-count iteration skips its V-cycle. Extension: with a relative tolerance it stops
// once ||r|| / ||b|| <= tol. Dot products accumulate in FP64 with a fixed grid and a fixed summation
// order (deterministic); scalars never leave the device, and a `done` flag gates every kernel of the
// loop so that iterations queued past convergence return at once.
//
// k_energy restates energyChange / chooseStep (solver.ts:430-534, newtonMultigrid.ts:582-603) in
// FP64: per node-owned term, the change of the incremental potential at alpha in {1, 1/2, 1/4, 1/8}
// by per-term differences and their magnitude sums, reduced deterministically, then the step rule.
#include <chrono>
#include <stdexcept>

#include "bd_internal.h"
#include "bounds.h"
#include "mg_dev.h"
#include "reduce.h"

namespace bd {

namespace {

struct PcgArgs {
  int N, Np;
  const uint8_t* fixed;
  Scal* sc;
  double* partials;
  float *r, *z, *p, *ap;
  double* delta;
  const float* rhs;
  float* zT;  // column-major copy of z (dual-layout fine level) or null
};

__global__ __launch_bounds__(1024) void k_pcg_start(PcgArgs A) {
  double v[1] = {0};
  const int Np = A.Np;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < A.N; n += gridDim.x * blockDim.x) {
    bool fx = A.fixed[n];
    for (int q = 0; q < 3; q++) {
      float b = fx ? 0.f : A.rhs[q * Np + n];
      A.r[q * Np + n] = b;
      A.z[q * Np + n] = 0.f;
      if (A.zT) A.zT[q * Np + n] = 0.f;
      A.delta[q * Np + n] = 0.0;
      A.p[q * Np + n] = 0.f;
      A.ap[q * Np + n] = 0.f;
      v[0] += (double)b * b;
    }
  }
  double tot[1];
  if (grid_sum<1>(v, A.partials, &A.sc->ticket, tot)) {
    Scal* s = A.sc;
    s->bb = tot[0];
    s->relres = tot[0] > 0 ? 1.0 : 0.0;
    s->iters = 0;
    s->rz = 0;
    s->beta = 0;
    s->need_true = 0;
    s->restart = 0;
    s->last = 0;
    s->replacements = 0;
    s->done = (s->tol > 0 && !(tot[0] > 0)) ? 1 : 0;
    s->learn_off = 0;
    s->warm = 1.0;
    for (int q = 0; q < 8; q++) s->hist[q] = 0;
  }
}

// rz = r . z ; beta = rz / rz_prev (first, or after a residual replacement: beta = 0). While the
// learned correction is on the preconditioner may change between iterations (the fallback switches
// it off), so beta is the flexible one of FCG(1): beta = -(z . A p_prev) / (p_prev . A p_prev), with
// A p_prev still in ap (Notay, Flexible conjugate gradients, SIAM J. Sci. Comput. 22, 2000).
__global__ __launch_bounds__(1024) void k_dot_rz(PcgArgs A, int first) {
  if (A.sc->done) return;
  double v[2] = {0, 0};
  const int Np = A.Np;
  const bool flex = A.sc->learn_on;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < A.N; n += gridDim.x * blockDim.x)
    for (int q = 0; q < 3; q++) {
      const double z = A.z[q * Np + n];
      v[0] += (double)A.r[q * Np + n] * z;
      if (flex) v[1] += z * (double)A.ap[q * Np + n];
    }
  double tot[2];
  if (grid_sum<2>(v, A.partials, &A.sc->ticket, tot)) {
    Scal* s = A.sc;
    if (first || s->restart)
      s->beta = 0.0;
    else
      s->beta = flex ? (s->pap > 0 ? -tot[1] / s->pap : 0.0) : tot[0] / s->rz;
    s->restart = 0;
    s->rz = tot[0];
  }
}

__global__ __launch_bounds__(256) void k_update_p(PcgArgs A) {
  if (A.sc->done) return;
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= A.N) return;
  const int Np = A.Np;
  float beta = (float)A.sc->beta;
  for (int q = 0; q < 3; q++) A.p[q * Np + n] = A.z[q * Np + n] + beta * A.p[q * Np + n];
}

__global__ __launch_bounds__(1024) void k_update_xr(PcgArgs A) {
  if (A.sc->done) return;
  const int Np = A.Np;
  const double alpha_d = A.sc->alpha;
  const float alpha = (float)alpha_d;
  double v[1] = {0};
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < A.N; n += gridDim.x * blockDim.x)
    for (int q = 0; q < 3; q++) {
      A.delta[q * Np + n] += alpha_d * (double)A.p[q * Np + n];
      float r = A.r[q * Np + n] - alpha * A.ap[q * Np + n];
      A.r[q * Np + n] = r;
      A.z[q * Np + n] = 0.f;
      if (A.zT) A.zT[q * Np + n] = 0.f;
      v[0] += (double)r * r;
    }
  double tot[1];
  if (grid_sum<1>(v, A.partials, &A.sc->ticket, tot)) {
    Scal* s = A.sc;
    s->rr = tot[0];
    s->relres = s->bb > 0 ? sqrt(tot[0] / s->bb) : 0.0;
    s->iters += 1;
    if (s->learn_on && !s->learn_off) {
      // fallback to plain MG-PCG when the residual has not dropped below its best for 16 iterations
      // (hist[0]: the best relative residual, hist[1]: its iteration)
      const int it = s->iters;
      if (it == 1 || s->relres < s->hist[0]) {
        s->hist[0] = s->relres;
        s->hist[1] = it;
      } else if (it - (int)s->hist[1] >= 16) {
        s->learn_off = 1;
        s->restart = 1;
        s->learn_fallbacks += 1;
      }
    }
    if (s->tol > 0) {
      // the recursive FP32 residual drifts from the true one: confirm with an FP64 recomputation
      s->last = s->iters >= s->max_iters;
      if (s->relres <= s->tol || s->last) s->need_true = 1;
    } else if (s->iters >= s->fixed_iters) {
      s->done = 1;
    }
  }
}

// Residual replacement (tolerance mode): r = b - A d with d in FP64 and FP64 accumulation over the
// FP32 operator; stop when the true relative residual meets the tolerance, otherwise continue from
// the true residual with a restarted search direction (iterative refinement of the FP32 inner solve).
__global__ __launch_bounds__(1024) void k_true_residual(PcgArgs A, const float* op, Geo geo) {
  if (A.sc->done || !A.sc->need_true) return;
  const int Np = A.Np;
  double v[1] = {0};
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < A.N; n += gridDim.x * blockDim.x) {
    double y[3] = {0, 0, 0};
    if (!A.fixed[n]) fine_apply_f64(op, Np, geo, A.delta, n, y);
    for (int q = 0; q < 3; q++) {
      double r = A.fixed[n] ? 0.0 : (double)A.rhs[q * Np + n] - y[q];
      A.r[q * Np + n] = (float)r;
      v[0] += r * r;
    }
  }
  double tot[1];
  if (grid_sum<1>(v, A.partials, &A.sc->ticket, tot)) {
    Scal* s = A.sc;
    s->rr = tot[0];
    s->relres = s->bb > 0 ? sqrt(tot[0] / s->bb) : 0.0;
    s->need_true = 0;
    s->replacements += 1;
    if (s->relres <= s->tol || s->last) {
      s->done = 1;
    } else {
      s->restart = 1;
    }
  }
}

// Ap = A p (fixed rows 0) and p . Ap ; alpha = rz / pAp, stop when pAp <= 0
__device__ __forceinline__ void blk_load6(const float* op, int Np, int base, int row, float o[6]) {
#pragma unroll
  for (int q = 0; q < 6; q++) o[q] = op[(base + q) * Np + row];
}

// Each block takes a contiguous chunk of nodes (a few lattice rows), so that the mirrored couplings a
// node reads from the rows below were fetched by the same block moments before (L2 hits); the loads
// of a node's 13 couplings are branch-free (clamped indices, masked products) and issue together.
__global__ __launch_bounds__(1024) void k_matvec_dot_fine(PcgArgs A, const float* op, Geo geo) {
  if (A.sc->done) return;
  const int Np = A.Np;
  double v[1] = {0};
  const int chunk = (A.N + gridDim.x - 1) / gridDim.x;
  const int n0 = blockIdx.x * chunk, n1 = min(A.N, n0 + chunk);
  const float* p = A.p;
  for (int n = n0 + threadIdx.x; n < n1; n += blockDim.x) {
    float y[3] = {0.f, 0.f, 0.f};
    const int l = geo_layer(geo, n);
    const int r0 = n - geo.off[l], nx = geo.nx[l], ny = geo.ny[l];
    const int j = r0 / nx, i = r0 - j * nx;
    auto acc = [&](bool in, int base, int src, int m) {
      float b[6], t[3];
      blk_load6(op, Np, base, in ? src : n, b);
      const int mm = in ? m : n;
      const float xm[3] = {p[mm], p[Np + mm], p[2 * Np + mm]};
      sym_mul(b, xm, t);
      const float w = in ? 1.f : 0.f;
      y[0] += w * t[0];
      y[1] += w * t[1];
      y[2] += w * t[2];
    };
    auto accs = [&](bool in, int plane, int src, int m) {
      const float sc = in ? op[plane * Np + src] : 0.f;
      const int mm = in ? m : n;
      y[0] += sc * p[mm];
      y[1] += sc * p[Np + mm];
      y[2] += sc * p[2 * Np + mm];
    };
    const bool e = i + 1 < nx, w = i > 0, no = j + 1 < ny, so = j > 0;
    acc(true, F_C, n, n);
    acc(e, F_E, n, n + 1);
    acc(w, F_E, n - 1, n - 1);
    acc(no, F_N, n, n + nx);
    acc(so, F_N, n - nx, n - nx);
    acc(e && no, F_NE, n, n + nx + 1);
    acc(w && so, F_NE, n - nx - 1, n - nx - 1);
    acc(w && no, F_NW, n, n + nx - 1);
    acc(e && so, F_NW, n - nx + 1, n - nx + 1);
    accs(i + 2 < nx, F_EE, n, n + 2);
    accs(i >= 2, F_EE, n - 2, n - 2);
    accs(j + 2 < ny, F_NN, n, n + 2 * nx);
    accs(j >= 2, F_NN, n - 2 * nx, n - 2 * nx);
    if (A.fixed[n]) y[0] = y[1] = y[2] = 0.f;
    for (int q = 0; q < 3; q++) {
      A.ap[q * Np + n] = y[q];
      v[0] += (double)p[q * Np + n] * (double)y[q];
    }
  }
  double tot[1];
  if (grid_sum<1>(v, A.partials, &A.sc->ticket, tot)) {
    Scal* s = A.sc;
    s->pap = tot[0];
    if (!(tot[0] > 0)) {
      s->done = 1;
      s->alpha = 0;
    } else {
      s->alpha = s->rz / tot[0];
    }
  }
}

// ---------------------------------------------------------------- line-search energy (FP64)
struct EnergyArgs {
  int N, Np;
  Geo geo;
  const DevLayer* layers;
  double h;
  const float *x, *dx, *pred, *mass;
  const double* d;
  const double* cn0;
  const uint8_t* pinned;
  const double *rest, *intact;
  const uint32_t* cut;
  const uint16_t* flags;
  const int* grip_of;
  const float* grip_w;
  const double* grip_off;
  const double* grip_goal_k;
  Scal* sc;
  double* partials;
  const float* ext;  // external terms (self-contact) or null
  int floor_on;
  double floor_z;
  // detached pieces' curved steps (7 planes: rotation w xyz, flag, arm r xyz), or null
  const float* curve;
};

__device__ __forceinline__ bool act(const EnergyArgs& E, int s) { return (E.flags[s] & (SEG_EXISTS | SEG_BROKEN)) == SEG_EXISTS; }

__device__ __forceinline__ void ld3d(const float* p, int Np, int n, double o[3]) {
  o[0] = p[n];
  o[1] = p[Np + n];
  o[2] = p[2 * Np + n];
}
__device__ __forceinline__ void ld3d(const double* p, int Np, int n, double o[3]) {
  o[0] = p[n];
  o[1] = p[Np + n];
  o[2] = p[2 * Np + n];
}

__device__ __forceinline__ double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// lenChange (solver.ts:447-451)
__device__ __forceinline__ void len_change(const double* e0, const double* de, double a, double l0, double& dl, double& la) {
  double ea[3] = {e0[0] + a * de[0], e0[1] + a * de[1], e0[2] + a * de[2]};
  la = sqrt(dot3(ea, ea));
  dl = (2 * a * dot3(e0, de) + a * a * dot3(de, de)) / fmax(la + l0, 1e-12);
}
// the same for a full displacement D of the term's vector (energy.wgsl:90-94)
__device__ __forceinline__ void len_change_d(const double* e0, const double* D, double l0, double& dl, double& la) {
  double ea[3] = {e0[0] + D[0], e0[1] + D[1], e0[2] + D[2]};
  la = sqrt(dot3(ea, ea));
  dl = (2 * dot3(e0, D) + dot3(D, D)) / fmax(la + l0, 1e-12);
}
__device__ __forceinline__ void symv6(const double* H, const double* v, double* o) {
  o[0] = H[0] * v[0] + H[1] * v[1] + H[2] * v[2];
  o[1] = H[1] * v[0] + H[3] * v[1] + H[4] * v[2];
  o[2] = H[2] * v[0] + H[4] * v[1] + H[5] * v[2];
}

// CURVE: some nodes belong to detached pieces and step along their piece's rotation
// (energy.wgsl:65-73): at step length a their step is a d + rotationCurve(a w, r).
template <bool CURVE>
__global__ __launch_bounds__(256) void k_energy(EnergyArgs E) {
  const double al[4] = {1.0, 0.5, 0.25, 0.125};
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // e[4], m[4]
  const int Np = E.Np;
  const double h = E.h;
  auto add = [&](int k, double v) {
    acc[k] += v;
    acc[4 + k] += fabs(v);
  };
  auto quad = [&](double gd, double q) {
    for (int k = 0; k < 4; k++) {
      acc[k] += al[k] * gd + 0.5 * al[k] * al[k] * q;
      acc[4 + k] += fabs(al[k] * gd) + fabs(0.5 * al[k] * al[k] * q);
    }
  };
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < E.N; n += gridDim.x * blockDim.x) {
    const int l = geo_layer(E.geo, n);
    const DevLayer& L = E.layers[l];
    const int nx = E.geo.nx[l], ny = E.geo.ny[l], r0 = n - E.geo.off[l];
    const int j = r0 / nx, i = r0 - j * nx;
    auto dir = [&](int m, double o[3]) {
      if (E.pinned[m]) {
        o[0] = o[1] = o[2] = 0;
      } else {
        ld3d(E.d, Np, m, o);
      }
    };
    // curved steps of detached pieces' nodes
    auto curved = [&](int m) { return CURVE && !E.pinned[m] && E.curve[3 * Np + m] >= 0.5f; };
    auto stepk = [&](int m, const double* dm, int k, double* o) {
      for (int q = 0; q < 3; q++) o[q] = al[k] * dm[q];
      if (curved(m)) {
        double th[3], r[3], c[3];
        for (int q = 0; q < 3; q++) {
          th[q] = al[k] * (double)E.curve[q * Np + m];
          r[q] = E.curve[(4 + q) * Np + m];
        }
        rotation_curve(th, r, c);
        for (int q = 0; q < 3; q++) o[q] += c[q];
      }
    };
    // change of a node-local quadratic term g . s + s^T H s / 2 over the node's curved steps
    auto quad_curved = [&](const double* g, const double* H, const double (*sk)[3]) {
      for (int k = 0; k < 4; k++) {
        double Hs[3];
        symv6(H, sk[k], Hs);
        double a = dot3(g, sk[k]), b = 0.5 * dot3(sk[k], Hs);
        acc[k] += a + b;
        acc[4 + k] += fabs(a) + fabs(b);
      }
    };
    double dn[3], d0[3], xn[3];
    dir(n, dn);
    ld3d(E.dx, Np, n, d0);
    ld3d(E.x, Np, n, xn);
    const bool cn = curved(n);
    double sn[4][3];
    if (cn)
      for (int k = 0; k < 4; k++) stepk(n, dn, k, sn[k]);
    if (!E.pinned[n]) {
      double w = (double)E.mass[n] / (h * h);
      double pr[3];
      ld3d(E.pred, Np, n, pr);
      if (!cn) {
        quad(w * (dn[0] * (d0[0] - pr[0]) + dn[1] * (d0[1] - pr[1]) + dn[2] * (d0[2] - pr[2])), w * dot3(dn, dn));
      } else {
        const double g[3] = {w * (d0[0] - pr[0]), w * (d0[1] - pr[1]), w * (d0[2] - pr[2])}, H[6] = {w, 0, 0, w, 0, w};
        quad_curved(g, H, sn);
      }
      int gi = E.grip_of[n];
      if (gi >= 0) {
        const double* G = E.grip_goal_k + 4 * gi;
        double kg = G[3] * (double)E.grip_w[n];
        double g[3];
        for (int q = 0; q < 3; q++) g[q] = kg * ((xn[q] - G[q]) + d0[q] - E.grip_off[q * Np + n]);
        if (kg > 0 && !cn) quad(dot3(g, dn), kg * dot3(dn, dn));
        if (kg > 0 && cn) {
          const double H[6] = {kg, 0, 0, kg, 0, kg};
          quad_curved(g, H, sn);
        }
      }
      // external terms (contact): quadratic model from their gradient and Hessian at this iterate
      // (energy.wgsl:106-110)
      if (E.ext) {
        double ge[3], He[6];
        for (int q = 0; q < 3; q++) ge[q] = E.ext[q * Np + n];
        for (int q = 0; q < 6; q++) He[q] = E.ext[(3 + q) * Np + n];
        double Hd0 = He[0] * dn[0] + He[1] * dn[1] + He[2] * dn[2];
        double Hd1 = He[1] * dn[0] + He[3] * dn[1] + He[4] * dn[2];
        double Hd2 = He[2] * dn[0] + He[4] * dn[1] + He[5] * dn[2];
        if (!cn)
          quad(dot3(ge, dn), dn[0] * Hd0 + dn[1] * Hd1 + dn[2] * Hd2);
        else
          quad_curved(ge, He, sn);
      }
      // floor contact: its quadratic model at this iterate (energy.wgsl:112-126)
      if (E.floor_on) {
        double reach = 0.5 * L.thickness + (double)kFloorBulge;
        double s0 = reach - (xn[2] + d0[2] - E.floor_z);
        if (s0 > 0) {
          double kc = kFloorK;
          double lam = kc * s0 * s0 / reach;
          double kf = (double)kFloorMu * lam / fmax(sqrt(d0[0] * d0[0] + d0[1] * d0[1]), (double)kFloorSpeed * h);
          if (!cn) {
            quad(kf * (d0[0] * dn[0] + d0[1] * dn[1]) - lam * dn[2], kf * (dn[0] * dn[0] + dn[1] * dn[1]) + (2.0 * kc * s0 / reach) * dn[2] * dn[2]);
          } else {
            const double g[3] = {kf * d0[0], kf * d0[1], -lam}, H[6] = {kf, 0, 0, kf, 0, 2.0 * kc * s0 / reach};
            quad_curved(g, H, sn);
          }
        }
      }
      if (L.air_drag > 0) {
        double k = (L.air_drag * L.pu * L.pv) / h;
        double t[2][3], ll[2];
        for (int f = 0; f < 2; f++) {
          int pv = f == 0 ? (i > 0 ? n - 1 : -1) : (j > 0 ? n - nx : -1);
          int nv = f == 0 ? (i + 1 < nx ? n + 1 : -1) : (j + 1 < ny ? n + nx : -1);
          int a = (pv >= 0 && act(E, 2 * pv + f)) ? pv : n;
          int b = (nv >= 0 && act(E, 2 * n + f)) ? nv : n;
          double xa[3], xb[3];
          ld3d(E.x, Np, a, xa);
          ld3d(E.x, Np, b, xb);
          for (int q = 0; q < 3; q++) t[f][q] = xb[q] - xa[q];
          ll[f] = sqrt(dot3(t[f], t[f]));
        }
        double e[3], kpar, kperp;
        bool done = false;
        if (ll[0] > 0 && ll[1] > 0) {
          double c[3] = {t[0][1] * t[1][2] - t[0][2] * t[1][1], t[0][2] * t[1][0] - t[0][0] * t[1][2], t[0][0] * t[1][1] - t[0][1] * t[1][0]};
          double lc = sqrt(dot3(c, c));
          if (lc > 1e-3 * ll[0] * ll[1]) {
            for (int q = 0; q < 3; q++) e[q] = c[q] / lc;
            kpar = k;
            kperp = 0;
            done = true;
          }
        }
        if (!done) {
          if (ll[0] > 0 || ll[1] > 0) {
            int f = ll[0] >= ll[1] ? 0 : 1;
            for (int q = 0; q < 3; q++) e[q] = t[f][q] / ll[f];
            kpar = 0;
            kperp = k;
          } else {
            e[0] = 1;
            e[1] = 0;
            e[2] = 0;
            kpar = k;
            kperp = k;
          }
        }
        double D[9];
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) D[3 * r + c] = (kpar - kperp) * e[r] * e[c] + (r == c ? kperp : 0);
        double t1[3], t2[3];
        for (int r = 0; r < 3; r++) {
          t1[r] = D[3 * r] * d0[0] + D[3 * r + 1] * d0[1] + D[3 * r + 2] * d0[2];
          t2[r] = D[3 * r] * dn[0] + D[3 * r + 1] * dn[1] + D[3 * r + 2] * dn[2];
        }
        if (!cn) {
          quad(dot3(t1, dn), dot3(dn, t2));
        } else {
          const double H[6] = {D[0], D[1], D[2], D[4], D[5], D[8]};
          quad_curved(t1, H, sn);
        }
      }
    }
    for (int f = 0; f < 2; f++) {
      const DevYarn& y = L.y[f];
      int next = f == 0 ? (i + 1 < nx ? n + 1 : -1) : (j + 1 < ny ? n + nx : -1);
      int prev = f == 0 ? (i > 0 ? n - 1 : -1) : (j > 0 ? n - nx : -1);
      int s = 2 * n + f;
      if (next >= 0 && act(E, s)) {
        double Lr = E.rest[s], it = E.intact[s];
        double xb[3], db[3], dd[3], en[3], ec[3], de[3];
        ld3d(E.x, Np, next, xb);
        ld3d(E.dx, Np, next, db);
        dir(next, dd);
        for (int q = 0; q < 3; q++) {
          en[q] = xb[q] - xn[q];
          ec[q] = en[q] + db[q] - d0[q];
          de[q] = dd[q] - dn[q];
        }
        double l0 = sqrt(dot3(ec, ec)), ln = sqrt(dot3(en, en));
        double T0 = it * axial_tension<double>(y, l0 / Lr - 1);
        double c = y.visc / (Lr * h);
        const bool cv = cn || curved(next);
        for (int k = 0; k < 4; k++) {
          double dl, la;
          if (!cv) {
            len_change(ec, de, al[k], l0, dl, la);
          } else {
            double sb[3], D[3];
            stepk(next, dd, k, sb);
            for (int q = 0; q < 3; q++) D[q] = sb[q] - (cn ? sn[k][q] : al[k] * dn[q]);
            len_change_d(ec, D, l0, dl, la);
          }
          double Ta = it * axial_tension<double>(y, la / Lr - 1);
          add(k, 0.5 * (T0 + Ta) * dl);
          add(k, 0.5 * c * dl * (la + l0 - 2 * ln));
        }
      }
      if (prev >= 0 && next >= 0 && act(E, 2 * prev + f) && act(E, s)) {
        int s0 = 2 * prev + f;
        double wa, wc, kperb, fold;
        bend_geometry<double>(E.rest[s0], E.rest[s], wa, wc, kperb, fold);
        double sc = bend_scale<double>(E.intact[s0], E.intact[s], E.cut[s0], E.cut[s], y.B);
        double kb = y.bend_rig * sc * kperb;
        double kd = (kb * y.bend_visc) / h;
        double xa[3], xc[3], da[3], dc[3], dp[3], dq[3], rn[3], r0[3], dr[3];
        ld3d(E.x, Np, prev, xa);
        ld3d(E.x, Np, next, xc);
        ld3d(E.dx, Np, prev, da);
        ld3d(E.dx, Np, next, dc);
        dir(prev, dp);
        dir(next, dq);
        for (int q = 0; q < 3; q++) {
          rn[q] = wa * (xa[q] - xn[q]) + wc * (xc[q] - xn[q]);
          r0[q] = rn[q] + wa * (da[q] - d0[q]) + wc * (dc[q] - d0[q]);
          dr[q] = wa * dp[q] - 2 * dn[q] + wc * dq[q];
        }
        double q0 = dot3(r0, r0), l0 = sqrt(q0), lrn = sqrt(dot3(rn, rn));
        const bool cv = cn || curved(prev) || curved(next);
        for (int k = 0; k < 4; k++) {
          double a = al[k];
          double dqa, dl, la;
          if (!cv) {
            dqa = 2 * a * dot3(r0, dr) + a * a * dot3(dr, dr);
            len_change(r0, dr, a, l0, dl, la);
          } else {
            double sa[3], sc2[3], sm[3], D[3];
            stepk(prev, dp, k, sa);
            stepk(next, dq, k, sc2);
            for (int q = 0; q < 3; q++) sm[q] = cn ? sn[k][q] : a * dn[q];
            for (int q = 0; q < 3; q++) D[q] = wa * sa[q] - 2 * sm[q] + wc * sc2[q];
            dqa = 2 * dot3(r0, D) + dot3(D, D);
            len_change_d(r0, D, l0, dl, la);
          }
          add(k, 0.5 * kb * bend_psi_delta(q0, q0 + dqa, dqa, fold));
          add(k, 0.5 * kd * dl * (la + l0 - 2 * lrn));
        }
      }
    }
    if (i + 1 < nx && j + 1 < ny) {
      int r = n + 1, t = n + nx, rt = r + nx;
      auto io = [&](int s) { return act(E, s) ? E.intact[s] : 0.0; };
      double ks = (L.shear_k * (double)E.cn0[n]) / L.cn0;
      double jam = shear_jam<double>(cell_area<double>(io(2 * n), io(2 * t), io(2 * n + 1), io(2 * r + 1)));
      double Ld = sqrt(E.rest[2 * n] * E.rest[2 * n] + E.rest[2 * n + 1] * E.rest[2 * n + 1]);
      bool cu = act(E, 2 * n), cv = act(E, 2 * n + 1), rv = act(E, 2 * r + 1), tu = act(E, 2 * t);
      for (int dg = 0; dg < 2; dg++) {
        bool on = dg == 0 ? ((cu && rv) || (cv && tu)) : ((cu && cv) || (rv && tu));
        if (!on) continue;
        int p = dg == 0 ? n : r, q2 = dg == 0 ? rt : t;
        double xp[3], xq[3], ddp[3], ddq[3], vp[3], vq[3], ec[3], de[3];
        ld3d(E.x, Np, p, xp);
        ld3d(E.x, Np, q2, xq);
        ld3d(E.dx, Np, p, ddp);
        ld3d(E.dx, Np, q2, ddq);
        dir(p, vp);
        dir(q2, vq);
        for (int z = 0; z < 3; z++) {
          ec[z] = (xq[z] - xp[z]) + (ddq[z] - ddp[z]);
          de[z] = vq[z] - vp[z];
        }
        double l0 = sqrt(dot3(ec, ec));
        double f0, k0, e0;
        shear_law<double>(ks, Ld, l0, jam, f0, k0, e0);
        const bool cv = curved(p) || curved(q2);
        for (int k = 0; k < 4; k++) {
          double dl, la, fa, ka, ea;
          if (!cv) {
            len_change(ec, de, al[k], l0, dl, la);
          } else {
            double sp[3], sq[3], D[3];
            stepk(p, vp, k, sp);
            stepk(q2, vq, k, sq);
            for (int z = 0; z < 3; z++) D[z] = sq[z] - sp[z];
            len_change_d(ec, D, l0, dl, la);
          }
          shear_law<double>(ks, Ld, la, jam, fa, ka, ea);
          add(k, 0.5 * (f0 + fa) * dl);
        }
      }
    }
  }
  double tot[8];
  if (grid_sum<8>(acc, E.partials, &E.sc->ticket, tot)) {
    Scal* s = E.sc;
    for (int k = 0; k < 4; k++) {
      s->e[k] = tot[k];
      s->m[k] = tot[4 + k];
    }
    // chooseStep (newtonMultigrid.ts:597-603)
    int best = 0;
    for (int k = 1; k < 4; k++)
      if (tot[k] < tot[best]) best = k;
    double alpha;
    bool all_up = true;
    for (int k = 0; k < 4; k++)
      if (!(tot[k] > kLineNoise * tot[4 + k])) all_up = false;
    if (tot[best] < -kLineNoise * tot[4 + best])
      alpha = al[best];
    else if (all_up)
      alpha = 0;
    else
      alpha = 1;
    if (alpha == 0) s->rejections += 1;
    s->line_alpha = alpha;
  }
}

__global__ void k_set_alpha(Scal* sc, double a) { sc->line_alpha = a; }

// the next k_true_residual runs: r = b - A delta from the rigid-mode start
__global__ void k_need_true(Scal* sc) {
  if (!sc->done) sc->need_true = 1;
}

__global__ void k_record(Scal* sc, int it) {
  sc->rec_iters[it] = sc->iters;
  sc->rec_relres[it] = sc->relres;
  sc->rec_alpha[it] = sc->line_alpha;
  sc->rec_warm[it] = sc->warm;
}

__global__ void k_reset_counters(Scal* sc) {
  for (int k = 0; k < BD_MAX_NEWTON; k++) {
    sc->max_update[k] = 0u;
    sc->rec_iters[k] = 0;
    sc->rec_relres[k] = 0;
    sc->rec_alpha[k] = 0;
    sc->rec_warm[k] = 1.0;
  }
}

// the learned warm start: delta = z (the correction applied to b), z cleared; the true residual follows
__global__ __launch_bounds__(256) void k_warm_apply(PcgArgs A) {
  if (A.sc->done) return;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n == 0) A.sc->need_true = 1;
  if (n >= A.N) return;
  const int Np = A.Np;
  for (int q = 0; q < 3; q++) {
    A.delta[q * Np + n] = A.z[q * Np + n];
    A.z[q * Np + n] = 0.f;
  }
}

// A warm start d0 = P Z y with y = (Z^T A_l Z)^-1 Z^T R b lowers the energy 1/2 d^T A d - b^T d by
// 1/2 c^T y >= 0 (the A-norm error never grows); its 2-norm residual may still exceed ||b||, so the
// start is only dropped when it is not finite or absurd (zero start, r = b, correction off).
__global__ __launch_bounds__(256) void k_warm_guard(PcgArgs A) {
  Scal* sc = A.sc;
  const bool bad = !(sc->relres < 1e3);
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (!bad) {
    if (n == 0) sc->warm = sc->relres;
    return;
  }
  if (n == 0) {
    // the fields gave no usable start (no decrease, or not finite): plain MG-PCG for this solve
    sc->warm = 1.0;
    sc->done = 0;
    sc->relres = 1.0;
    sc->restart = 1;
    sc->learn_off = 1;
    sc->learn_fallbacks += 1;
  }
  if (n >= A.N) return;
  const int Np = A.Np;
  for (int q = 0; q < 3; q++) {
    A.delta[q * Np + n] = 0.0;
    A.r[q * Np + n] = A.fixed[n] ? 0.f : A.rhs[q * Np + n];
  }
}

__global__ void k_pcg_setup(Scal* sc, double tol, int max_iters, int fixed_iters, int learn) {
  sc->learn_on = learn;
  sc->tol = tol;
  sc->max_iters = max_iters;
  sc->fixed_iters = fixed_iters;
  sc->done = 0;
  sc->ticket = 0u;
}

PcgArgs pcg_args(Solver& s) {
  Level& L = s.lv[0];
  return PcgArgs{s.N, s.Np, L.fixed, s.scal, s.partials, L.b, L.x, s.p, s.ap, s.delta, s.rhs, L.dual ? L.xT : nullptr};
}

int dot_grid(int N) {
  int g = (N + kDotThreads - 1) / kDotThreads;
  return g < kDotBlocks ? g : kDotBlocks;
}

}  // namespace

// PCG's convergence flag for the host: the flag, then (after a system-scope fence) the sequence
// number the host waits for, both in coherent pinned host memory
__global__ void k_post_flag(const Scal* sc, int* host, int seq) {
  volatile int* h = host;
  h[0] = sc->done;
  __threadfence_system();
  h[1] = seq;
}

// Waits for k_post_flag's sequence number by polling the pinned word (a stream synchronize would
// wake the host tens of microseconds after the copy ends), with a pause hint in the spin; after
// s.poll_timeout_ms without it, falls back to a stream synchronize and the device flag (the path a
// poll timeout of 0 takes at once).
bool wait_flag(Solver& s, int seq) {
  const auto t0 = std::chrono::steady_clock::now();
  const auto limit = std::chrono::milliseconds(s.poll_timeout_ms);
  for (unsigned spin = 0; s.poll_timeout_ms > 0; spin++) {
    if (__atomic_load_n(&s.flag_h[1], __ATOMIC_ACQUIRE) == seq) return __atomic_load_n(&s.flag_h[0], __ATOMIC_ACQUIRE) != 0;
    __builtin_ia32_pause();
    if ((spin & 1023) == 1023 && std::chrono::steady_clock::now() - t0 > limit) break;
  }
  s.poll_fallbacks += 1;
  BD_HIP(hipStreamSynchronize(s.stream));
  BD_HIP(hipMemcpy(&s.scal_h->done, &s.scal->done, sizeof(int), hipMemcpyDeviceToHost));
  return s.scal_h->done != 0;
}

// The solve's schedule into Scal for the one-workgroup path (the grid path sets its own below).
void launch_pcg_setup(Solver& s) {
  const bool tol_mode = s.P.pcg_rel_tol > 0;
  const int K = tol_mode ? std::max(1, s.P.pcg_max_iterations) : std::max(0, s.P.pcg_iterations);
  k_pcg_setup<<<1, 1, 0, s.stream>>>(s.scal, tol_mode ? s.P.pcg_rel_tol : 0.0, K, K, 0);
}

// One PCG solve on level 0 (newtonMultigrid.ts:537-576), leaving the direction in delta.
void pcg_solve(Solver& s) {
  if (s.pcg_mode == BD_PCG_WORKGROUP) {
    const std::string why = wg_unsupported(s);
    if (!why.empty()) throw std::runtime_error("no one-workgroup solve: " + why);
    launch_pcg_setup(s);
    launch_pcg_wg(s);
    return;
  }
  PcgArgs A = pcg_args(s);
  const bool tol_mode = s.P.pcg_rel_tol > 0;
  // captured into a graph: a fixed budget of iterations, the device's done flag gating those past
  // convergence, no host poll
  const bool graph = s.graph_pcg > 0;
  const int K = tol_mode ? (graph ? s.graph_pcg : std::max(1, s.P.pcg_max_iterations)) : std::max(0, s.P.pcg_iterations);
  const int G = dot_grid(s.N), B = grid_for(s.N, 256);
  const double N = s.N;
  const bool learned = s.learned != BD_LEARNED_OFF && learn_ready(s) && s.npieces == 0;
  const bool deflate = learned && s.learned == BD_LEARNED_DEFLATE;
  const bool record = !learned && tol_mode && learn_recording(s);
  k_pcg_setup<<<1, 1, 0, s.stream>>>(s.scal, tol_mode ? s.P.pcg_rel_tol : 0.0, K, K, learned ? 1 : 0);
  {
    ProfScope ps(s, PC_PCG_VECTOR, N * 85.0);
    k_pcg_start<<<G, kDotThreads, 0, s.stream>>>(A);
  }
  if (record) learn_record_begin(s, A.r);
  if (learned && K > 0) {
    // warm start d0 = P Z Ginv Z^T R b, kept only if it reduces the residual
    learn_correct(s, A.r, A.z);
    k_warm_apply<<<B, 256, 0, s.stream>>>(A);
    {
      ProfScope ps(s, PC_TRUE_RESIDUAL, 0.0);
      k_true_residual<<<G, kDotThreads, 0, s.stream>>>(A, s.lv[0].op, s.lv[0].geo);
    }
    k_warm_guard<<<B, 256, 0, s.stream>>>(A);
  }
  if (s.npieces > 0) {
    // detached pieces (rigid_apply.wgsl): start from each piece's exact rigid response x0 and the
    // residual b - A x0, so the direction is never worse than that start
    launch_rigid_start(s);
    k_need_true<<<1, 1, 0, s.stream>>>(s.scal);
    ProfScope ps(s, PC_TRUE_RESIDUAL, 0.0);
    k_true_residual<<<G, kDotThreads, 0, s.stream>>>(A, s.lv[0].op, s.lv[0].geo);
  }
  if (K == 0) return;
  vcycle(s, 0);
  if (deflate) learn_deflate(s, A.z);
  {
    ProfScope ps(s, PC_PCG_VECTOR, N * 24.0);
    k_dot_rz<<<G, kDotThreads, 0, s.stream>>>(A, 1);
  }
  {
    ProfScope ps(s, PC_PCG_VECTOR, N * 36.0);
    k_update_p<<<B, 256, 0, s.stream>>>(A);
  }
  for (int it = 0; it < K; it++) {
    {
      // fine row (58 floats) + p read + Ap write
      ProfScope ps(s, PC_MATVEC, N * 256.0);
      k_matvec_dot_fine<<<G, kDotThreads, 0, s.stream>>>(A, s.lv[0].op, s.lv[0].geo);
    }
    {
      ProfScope ps(s, PC_PCG_VECTOR, N * 108.0);
      k_update_xr<<<G, kDotThreads, 0, s.stream>>>(A);
    }
    if (tol_mode) {
      ProfScope ps(s, PC_TRUE_RESIDUAL, 0.0);
      k_true_residual<<<G, kDotThreads, 0, s.stream>>>(A, s.lv[0].op, s.lv[0].geo);
    }
    if (record && it + 1 == BD_NET_SLOW_ITER) learn_record_mid(s, A.r, A.delta);
    if (it + 1 == K) break;
    if (tol_mode && !graph) {
      // the device decides convergence; check it before queueing the next V-cycle
      const int seq = ++s.flag_seq;
      k_post_flag<<<1, 1, 0, s.stream>>>(s.scal, s.flag_d, seq);
      BD_HIP(hipGetLastError());
      if (wait_flag(s, seq)) break;
    }
    vcycle(s, 0);
    if (deflate) learn_deflate(s, A.z);
    {
      ProfScope ps(s, PC_PCG_VECTOR, N * 24.0);
      k_dot_rz<<<G, kDotThreads, 0, s.stream>>>(A, 0);
    }
    {
      ProfScope ps(s, PC_PCG_VECTOR, N * 36.0);
      k_update_p<<<B, 256, 0, s.stream>>>(A);
    }
  }
  if (record) learn_record_end(s, A.delta, A.rhs);
}

void launch_record(Solver& s, int it) { k_record<<<1, 1, 0, s.stream>>>(s.scal, it); }

__global__ void k_open_gate(Scal* s) { s->done = 0; }

// Clears the gate of the V-cycle kernels with a kernel on the solver's stream (ordered like every
// other launch, and capturable), so that a V-cycle, sweep or residual run outside PCG does its work.
void launch_open_gate(Solver& s) { k_open_gate<<<1, 1, 0, s.stream>>>(s.scal); }

void launch_reset_counters(Solver& s) { k_reset_counters<<<1, 1, 0, s.stream>>>(s.scal); }

void launch_energy(Solver& s) {
  if (!s.P.line_search) {
    k_set_alpha<<<1, 1, 0, s.stream>>>(s.scal, 1.0);
    return;
  }
  EnergyArgs E{s.N, s.Np, s.geo, s.layers, s.P.h, s.x, s.dx, s.pred, s.mass, s.delta, s.cn0, s.pinned, s.rest, s.intact,
               s.cut_mask, s.flags, s.grip_of, s.grip_w, s.grip_off, s.grip_goal_k, s.scal, s.partials,
               (s.contact_mode != kContactNone || s.grips_active) ? s.ext : nullptr, s.floor_on ? 1 : 0, (double)s.floor_z,
               s.npieces > 0 ? s.curve : nullptr};
  ProfScope ps(s, PC_ENERGY, s.N * 161.0);
  // 256-thread blocks keep the FP64 energy terms in registers
  if (E.curve)
    k_energy<true><<<std::min(kDotBlocks, grid_for(s.N, 256)), 256, 0, s.stream>>>(E);
  else
    k_energy<false><<<std::min(kDotBlocks, grid_for(s.N, 256)), 256, 0, s.stream>>>(E);
}

}  // namespace bd

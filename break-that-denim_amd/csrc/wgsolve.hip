// One Newton system's whole PCG solve inside one workgroup (BD_PCG_WORKGROUP), and many systems in
// one launch (a batch of patches: one workgroup each).
//
// Small systems (a 64 x 64 sheet, the patches of a batch) are bound by the latency of dependent
// launches: the grid path issues some fifty kernels per PCG iteration (a V-cycle's colour passes,
// residuals and transfers, the matvec and dot products) and each costs 5-10 us whatever its size.
// Here one workgroup of 1024 threads runs the same algorithm end to end with workgroup barriers
// between its phases: the start, and per iteration the matvec with p.Ap, the update with ||r||, the
// FP64 true residual when the recursive one meets the tolerance (residual replacement), the V-cycle
// and the flexible-free beta. Every phase is the grid path's own device code (the row gather and the
// block PCR line solve of smooth_dev.h, the residual row, restriction and prolongation of mg_dev.h,
// the FP64 true residual row), so the solve is the same preconditioned CG; only the FP64 sums run in
// another order (a workgroup tree instead of the grid's block partials).
//
// Line solves: a colour pass takes its lines in rounds; in a round every line of the colour gets
// R2 = ceil(len / 2) threads (two rows each, as the grid kernels' threads) and its own LDS records,
// up to kWgSlots threads in lines. Threads without a line run the same solve on a spare record with
// identity rows, so every thread meets every barrier of the reduction (the level count depends only
// on the line length, equal for all lines of one family on a one-layer level).
//
// Limits (checked by the host): one layer, FP32 operators, lines of at most kWgMaxLine nodes, no
// detached pieces and no learned path. The coarsest level is swept (BD_COARSE_SWEEPS) or solved by
// its dense FP64 inverse, one wave per row (BD_COARSE_EXACT; 3n x 3n doubles read per V-cycle).
#include <stdexcept>

#include "bd_internal.h"
#include "mg_dev.h"
#include "reduce.h"
#include "smooth_dev.h"

namespace bd {

namespace {

// 512 threads: the phases' registers (the two-row gather and the line solve's blocks) fit without
// spilling (a 1024-thread workgroup would cap each thread at 128 VGPRs)
constexpr int kWgThreads = 512;
constexpr int kWgSlots = 512;

struct WgLevel {
  const float* op;
  uint8_t* fixed;
  int N, Np, fine;
  Geo geo;
  float *x, *b, *r;
  const float* pw;  // prolongation weights to the next level (null on the coarsest)
};

struct WgSys {
  WgLevel lv[kMaxLevels];
  // each level's pass and transfer arguments, read in place from device memory (a local copy of
  // their lattice tables, indexed by layer, would live in scratch)
  SmoothArgs sa[kMaxLevels];
  XferArgs xf[kMaxLevels];
  int nlev, ncol, smooth_sweeps, coarse_sweeps, exact, cn;
  const double* cinv;
  Scal* sc;
  float *p, *ap;
  double* delta;
  const float* rhs;
};

// deterministic workgroup sum of one double per thread: waves by shuffles, then the waves in order;
// every thread gets the total
__device__ __forceinline__ double wg_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < kWgThreads / 64; w++) s += red[w];
  __syncthreads();
  return s;
}

// one colour pass: lines c, c + ncol, ... of family FAM (0: u-lines along i, 1: v-lines along j)
template <bool FINE, int FAM>
__device__ void wg_colour(const WgLevel& L, const SmoothArgs& S, int c, int ncol, float* cr) {
  const int nx = L.geo.nx[0], ny = L.geo.ny[0], off = L.geo.off[0];
  const int len = FAM == 0 ? nx : ny, lines = FAM == 0 ? ny : nx;
  if (c >= lines) return;
  const int count = (lines - c + ncol - 1) / ncol;
  const int R2 = (len + 1) >> 1;
  const int per = kWgSlots / R2;
  const OpF32 rd{L.op, L.Np};
  // u-lines: a line's rows on consecutive threads (rows 2t, 2t + 1 sit next to each other in memory);
  // v-lines: consecutive threads on consecutive lines of the colour (row t of neighbouring lines is
  // ncol nodes apart, where a thread walking down one v-line would stride nx nodes)
  const int tid = threadIdx.x;
  int li, t;
  if (FAM == 0 || tid >= per * R2) {
    li = tid / R2;
    t = tid - li * R2;
  } else {
    t = tid / per;
    li = tid - t * per;
  }
  for (int base = 0; base < count; base += per) {
    const int line = base + li;
    const bool active = li < per && line < count;
    // this thread's line; a thread without one works on line c's rows in the spare record and
    // stores nothing
    LineGeo G;
    const int idx = c + (active ? line : 0) * ncol;
    G.nx = nx;
    G.ny = ny;
    G.off = off;
    G.i0 = FAM == 0 ? 0 : idx;
    G.j0 = FAM == 0 ? idx : 0;
    G.ln = FAM == 0 ? make_int4(off + idx * nx, 1, nx, 0) : make_int4(off + idx, nx, ny, 1);
    const int ka = 2 * t, kb = 2 * t + 1;
    int na, nba, noa, ia, ja, nb, nbb, nob, ib, jb;
    row_index<FAM, false>(G, ka, na, nba, noa, ia, ja);
    row_index<FAM, false>(G, kb, nb, nbb, nob, ib, jb);
    float Da[6], Ua[9], ra[3], Db[6], Ub[9], rb[3];
    bool fa = true, fb = true;
    if (active && ka < len)
      gather_row<FINE, FAM, false>(S, rd, na, nba, ia, ja, nx, ny, Da, Ua, ra, fa);
    else
      set_identity(Da, Ua, ra);
    if (active && kb < len)
      gather_row<FINE, FAM, false>(S, rd, nb, nbb, ib, jb, nx, ny, Db, Ub, rb, fb);
    else
      set_identity(Db, Ub, rb);
    float xa[3], xb[3];
    float* mine = cr + (size_t)(active ? li : per) * kSolveFloats * R2;
    line_solve(mine, R2, t, len, Da, Ua, ra, Db, Ub, rb, xa, xb);
    if (active && ka < len && !fa) store_x(S, na, noa, xa, false);
    if (active && kb < len && !fb) store_x(S, nb, nob, xb, false);
    __syncthreads();
  }
}

// a sweep over every colour of level L: v-line colours first, then u-line colours (reversed for the
// post-smoothing sweep), as the grid path's launch_smooth
__device__ void wg_smooth(const WgLevel& L, const SmoothArgs& S, int ncol, int reverse, float* cr) {
  const int nc = 2 * ncol;
  for (int q = 0; q < nc; q++) {
    const int c = reverse ? nc - 1 - q : q;
    const int fam = c < ncol ? 1 : 0, cc = c < ncol ? c : c - ncol;
    if (L.fine) {
      if (fam) wg_colour<true, 1>(L, S, cc, ncol, cr);
      else wg_colour<true, 0>(L, S, cc, ncol, cr);
    } else {
      if (fam) wg_colour<false, 1>(L, S, cc, ncol, cr);
      else wg_colour<false, 0>(L, S, cc, ncol, cr);
    }
  }
}

// r = b - A x on level L (fixed rows 0)
__device__ void wg_residual(const WgLevel& L) {
  const OpF32 rd{L.op, L.Np};
  const int Np = L.Np;
  for (int n = threadIdx.x; n < L.N; n += kWgThreads) {
    if (L.fixed[n]) {
      L.r[n] = L.r[Np + n] = L.r[2 * Np + n] = 0.f;
      continue;
    }
    float y[3];
    if (L.fine) vc_apply_row<true>(rd, L.geo, L.x, Np, n, y);
    else vc_apply_row<false>(rd, L.geo, L.x, Np, n, y);
    for (int q = 0; q < 3; q++) L.r[q * Np + n] = L.b[q * Np + n] - y[q];
  }
}

// x = A_c^-1 b on the coarsest level: one wave per DOF row, FP64 products and sums
__device__ void wg_coarse_apply(const WgSys& Y, const WgLevel& L) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, n = Y.cn;
  for (int row = wid; row < n; row += kWgThreads / 64) {
    const double* r = Y.cinv + (size_t)row * n;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) {
      const int node = j / 3, c = j - 3 * node;
      s += r[j] * (double)L.b[c * L.Np + node];
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) {
      const int node = row / 3, c = row - 3 * node;
      L.x[c * L.Np + node] = (float)s;
    }
  }
}

// the symmetric V-cycle of the grid path (vcycle in multigrid.hip), one phase after another
__device__ void wg_vcycle(const WgSys& Y, float* cr) {
  const int last = Y.nlev - 1;
  for (int k = 0; k < last; k++) {
    const WgLevel& L = Y.lv[k];
    for (int sw = 0; sw < Y.smooth_sweeps; sw++) wg_smooth(L, Y.sa[k], Y.ncol, 0, cr);
    wg_residual(L);
    __syncthreads();
    const WgLevel& Cl = Y.lv[k + 1];
    const XferArgs& X = Y.xf[k];
    for (int I = threadIdx.x; I < Cl.N; I += kWgThreads) {
      float s[3];
      restrict_node(X, I, s);
      for (int q = 0; q < 3; q++) {
        Cl.b[q * Cl.Np + I] = s[q];
        Cl.x[q * Cl.Np + I] = 0.f;
      }
    }
    __syncthreads();
  }
  const WgLevel& C = Y.lv[last];
  if (Y.exact) wg_coarse_apply(Y, C);
  else
    for (int sw = 0; sw < Y.coarse_sweeps; sw++) wg_smooth(C, Y.sa[last], Y.ncol, sw & 1, cr);
  __syncthreads();
  for (int k = last - 1; k >= 0; k--) {
    const WgLevel& L = Y.lv[k];
    const XferArgs& X = Y.xf[k];
    for (int f = threadIdx.x; f < L.N; f += kWgThreads) {
      if (L.fixed[f]) continue;
      int l, i, j;
      node_ij(L.geo, f, l, i, j);
      float s[3];
      prolong_sum(X, f, l, i, j, s);
      for (int q = 0; q < 3; q++) L.x[q * L.Np + f] += s[q];
    }
    __syncthreads();
    for (int sw = 0; sw < Y.smooth_sweeps; sw++) wg_smooth(L, Y.sa[k], Y.ncol, 1, cr);
  }
}

// PCG on level 0 of system blockIdx.x (newtonMultigrid.ts:537-576 with the grid path's extensions:
// relative tolerance with FP64 residual replacement); k_pcg_setup has set the schedule in Scal
__global__ __launch_bounds__(kWgThreads) void k_pcg_wg(const WgSys* systems) {
  const WgSys& Y = systems[blockIdx.x];
  extern __shared__ __align__(16) float cr[];
  __shared__ double red[kWgThreads / 64];
  Scal* sc = Y.sc;
  const WgLevel& L0 = Y.lv[0];
  const int N = L0.N, Np = L0.Np;
  float* r = L0.b;  // the V-cycle's level-0 rhs is the residual, its solution z
  float* z = L0.x;
  const double tol = sc->tol;
  const bool tol_mode = tol > 0;
  const int K = sc->max_iters;
  // start: r = b (fixed rows 0), d = z = p = Ap = 0
  double v = 0.0;
  for (int n = threadIdx.x; n < N; n += kWgThreads) {
    const bool fx = L0.fixed[n];
    for (int q = 0; q < 3; q++) {
      const float b = fx ? 0.f : Y.rhs[q * Np + n];
      r[q * Np + n] = b;
      z[q * Np + n] = 0.f;
      Y.delta[q * Np + n] = 0.0;
      Y.p[q * Np + n] = 0.f;
      Y.ap[q * Np + n] = 0.f;
      v += (double)b * b;
    }
  }
  const double bb = wg_sum(v, red);
  double relres = bb > 0 ? 1.0 : 0.0, rz = 0.0, beta = 0.0, pap = 0.0, alpha = 0.0, rr = bb;
  int iters = 0, replacements = 0;
  bool done = tol_mode && !(bb > 0), restart = false, last = false;
  if (K > 0 && !done) {
    wg_vcycle(Y, cr);
    v = 0.0;
    for (int n = threadIdx.x; n < N; n += kWgThreads)
      for (int q = 0; q < 3; q++) v += (double)r[q * Np + n] * (double)z[q * Np + n];
    rz = wg_sum(v, red);
    for (int n = threadIdx.x; n < N; n += kWgThreads)
      for (int q = 0; q < 3; q++) Y.p[q * Np + n] = z[q * Np + n];
    __syncthreads();
    const OpF32 rd{L0.op, Np};
    for (int it = 0; it < K; it++) {
      // Ap and p.Ap
      v = 0.0;
      for (int n = threadIdx.x; n < N; n += kWgThreads) {
        float y[3] = {0.f, 0.f, 0.f};
        if (!L0.fixed[n]) vc_apply_row<true>(rd, L0.geo, Y.p, Np, n, y);
        for (int q = 0; q < 3; q++) {
          Y.ap[q * Np + n] = y[q];
          v += (double)Y.p[q * Np + n] * (double)y[q];
        }
      }
      pap = wg_sum(v, red);
      if (!(pap > 0)) {
        done = true;
        alpha = 0.0;
        break;
      }
      alpha = rz / pap;
      // d += alpha p, r -= alpha Ap, z = 0, ||r||
      const float af = (float)alpha;
      v = 0.0;
      for (int n = threadIdx.x; n < N; n += kWgThreads)
        for (int q = 0; q < 3; q++) {
          Y.delta[q * Np + n] += alpha * (double)Y.p[q * Np + n];
          const float rn = r[q * Np + n] - af * Y.ap[q * Np + n];
          r[q * Np + n] = rn;
          z[q * Np + n] = 0.f;
          v += (double)rn * rn;
        }
      rr = wg_sum(v, red);
      relres = bb > 0 ? sqrt(rr / bb) : 0.0;
      iters += 1;
      bool need_true = false;
      if (tol_mode) {
        last = iters >= K;
        need_true = relres <= tol || last;
      } else if (iters >= sc->fixed_iters) {
        done = true;
      }
      if (need_true) {
        // the FP64 true residual of the FP32 iterate: stop at the tolerance, else restart from it
        v = 0.0;
        for (int n = threadIdx.x; n < N; n += kWgThreads) {
          double y[3] = {0, 0, 0};
          if (!L0.fixed[n]) fine_apply_f64(L0.op, Np, L0.geo, Y.delta, n, y);
          for (int q = 0; q < 3; q++) {
            const double rt = L0.fixed[n] ? 0.0 : (double)Y.rhs[q * Np + n] - y[q];
            r[q * Np + n] = (float)rt;
            v += rt * rt;
          }
        }
        rr = wg_sum(v, red);
        relres = bb > 0 ? sqrt(rr / bb) : 0.0;
        replacements += 1;
        if (relres <= tol || last) done = true;
        else restart = true;
      }
      if (done || it + 1 == K) break;
      wg_vcycle(Y, cr);
      v = 0.0;
      for (int n = threadIdx.x; n < N; n += kWgThreads)
        for (int q = 0; q < 3; q++) v += (double)r[q * Np + n] * (double)z[q * Np + n];
      const double rz_new = wg_sum(v, red);
      beta = restart ? 0.0 : rz_new / rz;
      restart = false;
      rz = rz_new;
      const float bf = (float)beta;
      for (int n = threadIdx.x; n < N; n += kWgThreads)
        for (int q = 0; q < 3; q++) Y.p[q * Np + n] = z[q * Np + n] + bf * Y.p[q * Np + n];
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    sc->bb = bb;
    sc->rz = rz;
    sc->pap = pap;
    sc->alpha = alpha;
    sc->beta = beta;
    sc->rr = rr;
    sc->relres = relres;
    sc->iters = iters;
    sc->replacements = replacements;
    sc->need_true = 0;
    sc->restart = 0;
    sc->last = last ? 1 : 0;
    sc->done = 1;
    sc->learn_off = 0;
    sc->warm = 1.0;
  }
}

}  // namespace

// Why a solver cannot take the one-workgroup path ("" when it can).
std::string wg_unsupported(const Solver& s) {
  if (s.nl != 1) return "it has several layers";
  if (s.npieces > 0) return "detached pieces need the grid path's rigid start";
  if (s.learned != BD_LEARNED_OFF) return "the learned path is on";
  if (learn_recording(s)) return "sample recording is on";
  for (int k = 0; k < s.nlev; k++) {
    const Level& L = s.lv[k];
    if (L.oph) return "its coarse levels hold FP16 operators";
    if (L.ncol != kLineColors) return "a level smooths with fewer than 3 line colours";
    if (std::max(L.geo.nx[0], L.geo.ny[0]) > kWgMaxLine) return "its lines are longer than 512 nodes";
  }
  return "";
}

// The system's descriptor on the device (pointers fixed at creation) and the kernel's LDS.
void wg_prepare(Solver& s) {
  if (s.wg_sys) return;
  WgSys Y{};
  Y.nlev = s.nlev;
  Y.ncol = s.lv[0].ncol;
  Y.smooth_sweeps = std::max(1, s.P.smooth_sweeps);
  Y.coarse_sweeps = s.P.coarse_sweeps;
  Y.exact = coarse_exact(s) ? 1 : 0;
  Y.cn = s.cn;
  Y.cinv = s.cinv;
  Y.sc = s.scal;
  Y.p = s.p;
  Y.ap = s.ap;
  Y.delta = s.delta;
  Y.rhs = s.rhs;
  int maxr2 = 1;
  for (int k = 0; k < s.nlev; k++) {
    const Level& L = s.lv[k];
    Y.lv[k] = WgLevel{L.op, L.fixed, L.N, L.Np, L.fine ? 1 : 0, L.geo, L.x, L.b, L.r, k + 1 < s.nlev ? L.pw : nullptr};
    Y.sa[k] = SmoothArgs{L.op, L.fixed, L.Np, nullptr, L.x, L.b, nullptr, nullptr, L.geo, L.x};
    if (k + 1 < s.nlev) {
      const Level& C = s.lv[k + 1];
      Y.xf[k] = XferArgs{LvArgs{L.op, L.fixed, L.N, L.Np, L.geo}, C.geo, C.N, C.Np, L.pw, C.fixed, L.r, C.b, C.x, L.x,
                         nullptr, nullptr, nullptr};
    }
    maxr2 = std::max(maxr2, (std::max(L.geo.nx[0], L.geo.ny[0]) + 1) / 2);
  }
  s.wg_lds = sizeof(float) * (size_t)kSolveFloats * (kWgSlots + maxr2);
  // the kernel's limit is set once for its largest case, so that launches of other systems' sizes
  // (a batch's) stay within it
  static const size_t kMaxWgLds = sizeof(float) * (size_t)kSolveFloats * (kWgSlots + (kWgMaxLine + 1) / 2);
  static_assert(sizeof(float) * (size_t)kSolveFloats * (kWgSlots + (kWgMaxLine + 1) / 2) + 1024 <= 160 * 1024, "LDS");
  BD_HIP(hipFuncSetAttribute((const void*)k_pcg_wg, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxWgLds));
  void* d = nullptr;
  BD_HIP(hipMalloc(&d, sizeof(WgSys)));
  BD_HIP(hipMemcpy(d, &Y, sizeof(WgSys), hipMemcpyHostToDevice));
  s.wg_sys = d;
}

void wg_free(Solver& s) {
  if (s.wg_sys) (void)hipFree(s.wg_sys);
  s.wg_sys = nullptr;
}

// One solve (k_pcg_setup has run on the stream).
void launch_pcg_wg(Solver& s) {
  coarse_join(s);
  ProfScope ps(s, PC_PCG_VECTOR, 0.0);
  k_pcg_wg<<<1, kWgThreads, s.wg_lds, s.stream>>>(static_cast<const WgSys*>(s.wg_sys));
}

size_t wg_sys_bytes() { return sizeof(WgSys); }

// Many systems' solves in one launch: `systems` holds `count` descriptors back to back (copies of
// each solver's wg_sys), lds the largest of their wg_lds.
void launch_pcg_wg_many(const void* systems, int count, size_t lds, hipStream_t st) {
  k_pcg_wg<<<count, kWgThreads, lds, st>>>(static_cast<const WgSys*>(systems));
}

}  // namespace bd

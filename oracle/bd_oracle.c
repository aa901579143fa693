/*
 * bd_oracle.c -- FP64 CPU oracle of the implicit cloth solve.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, the smoke check and bench.py's cpu_baseline leg load this library, and only as the
 * checker (or the timed CPU baseline). The product (libbd.so) never links or calls it.
 *
 * A plain-C restatement of the reference's f64 solver, followed function by function:
 *   laws ........ src/simulation/axial.ts:16-85, src/simulation/laws.ts:18-159
 *   solver ...... src/simulation/reference/solver.ts:41-749 (ReferenceSolver)
 *   multigrid ... src/simulation/reference/newtonMultigrid.ts:38-733
 *   lattice ..... src/simulation/lattice.ts:13-66
 *   yarn field .. src/simulation/yarnField.ts:68-499
 * Extensions beyond the reference, all off unless asked for: several grips (the reference holds one),
 * PCG to a relative residual (the reference runs a fixed count), gated extra Newton iterations
 * (gate.wgsl semantics). Compile with -ffp-contract=off so products are rounded as written.
 */
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/bd.h"

#define SEG_EXISTS 1
#define SEG_BROKEN 2
#define SEG_SLIPPED 8
#define DAMAGE_CUT 255
#define DAMAGE_PULLED 254
#define DAMAGE_SEPARATED 254
#define AREA_MIN 0.02
#define STENCIL 25
#define CENTER 12
#define LINK_MIN_INTACT 0.02
#define PIVOT_SHIFT 2e-6
#define LINE_NOISE 4e-6
#define BEND_S0 0.85
#define SHEAR_JAM 0.16
#define SHEAR_RAMP 0.08
#define SHEAR_LOCK 40.0
#define SHARE_FRACTION 0.6
#define SHARE_MAXRUN 6
#define SHARE_REACH 1.2e-3
#define RECOIL_MAX 12
#define MAXLV 16

static const double LINE_ALPHAS[4] = {1.0, 0.5, 0.25, 0.125};

static inline int offset_index(int di, int dj) { return (dj + 2) * 5 + (di + 2); }

/* ---------------------------------------------------------------- laws (axial.ts, laws.ts) */

double or_axial_tension(const bd_yarn* p, double eps) {
  if (eps <= 0) return p->ea_compress * eps;
  double ec = p->crimp_strain, dk = p->ea_high - p->ea_low;
  if (eps <= ec) return p->ea_low * eps + (dk * eps * eps) / (2 * ec);
  return p->ea_low * ec + (dk * ec) / 2 + p->ea_high * (eps - ec);
}

double or_axial_stiffness(const bd_yarn* p, double eps) {
  if (eps <= 0) return p->ea_compress;
  double ec = p->crimp_strain;
  if (eps <= ec) return p->ea_low + ((p->ea_high - p->ea_low) * eps) / ec;
  return p->ea_high;
}

double or_axial_energy_density(const bd_yarn* p, double eps) {
  if (eps <= 0) return 0.5 * p->ea_compress * eps * eps;
  double ec = p->crimp_strain, dk = p->ea_high - p->ea_low;
  if (eps <= ec) return 0.5 * p->ea_low * eps * eps + (dk * eps * eps * eps) / (6 * ec);
  double tc = p->ea_low * ec + (dk * ec) / 2;
  double ece = 0.5 * p->ea_low * ec * ec + (dk * ec * ec) / 6;
  double d = eps - ec;
  return ece + tc * d + 0.5 * p->ea_high * d * d;
}

double or_strain_at_tension(const bd_yarn* p, double t) {
  double ec = p->crimp_strain, dk = p->ea_high - p->ea_low;
  double tc = p->ea_low * ec + (dk * ec) / 2;
  if (t >= tc) return ec + (t - tc) / p->ea_high;
  double a = dk / (2 * ec);
  return (2 * t) / (p->ea_low + sqrt(p->ea_low * p->ea_low + 4 * a * t));
}

double or_yarn_damage(const bd_yarn* p, double yarns, double area, double peak, double soft, double eps) {
  double ep = or_strain_at_tension(p, (peak * yarns) / area);
  if (eps <= ep) return 0;
  if (eps >= ep + soft) return 1;
  double tenv = peak * (1 - (eps - ep) / soft);
  double tel = (area * or_axial_tension(p, eps)) / yarns;
  double d = 1 - tenv / fmax(tel, 1e-12);
  return fmin(1, fmax(0, d));
}

void or_bend_psi(double q, double a, double out[3]) {
  double s = q / a;
  if (s < BEND_S0) {
    double d = 1 - s;
    out[0] = q / d;
    out[1] = 1 / (d * d);
    out[2] = 2 / (a * d * d * d);
    return;
  }
  double d0 = 1 - BEND_S0, q0 = BEND_S0 * a;
  double p1 = 1 / (d0 * d0), p2 = 2 / (a * d0 * d0 * d0), t = q - q0;
  out[0] = q0 / d0 + p1 * t + 0.5 * p2 * t * t;
  out[1] = p1 + p2 * t;
  out[2] = p2;
}

double or_bend_psi_delta(double q0, double q1, double dq, double a) {
  double s0 = q0 / a, s1 = q1 / a;
  if (s0 < BEND_S0 && s1 < BEND_S0) return dq / ((1 - s0) * (1 - s1));
  double qb = BEND_S0 * a, d0 = 1 - BEND_S0;
  double p1 = 1 / (d0 * d0), p2 = 2 / (a * d0 * d0 * d0);
  if (s0 >= BEND_S0 && s1 >= BEND_S0) return dq * (p1 + 0.5 * p2 * (q0 - qb + (q1 - qb)));
  double lo = fmin(q0, q1);
  double t = fmax(q0, q1) - qb;
  double up = (qb - lo) / ((1 - lo / a) * d0) + t * (p1 + 0.5 * p2 * t);
  return q1 >= q0 ? up : -up;
}

/* bendGeometry (laws.ts:52-56): out = [wa, wc, kPerB, fold] */
void or_bend_geometry(double La, double Lc, double out[4]) {
  double s = La + Lc;
  double lh = (2 * La * Lc) / s;
  out[0] = (2 * Lc) / s;
  out[1] = (2 * La) / s;
  out[2] = 1 / (lh * lh * 0.5 * s);
  out[3] = 4 * lh * lh;
}

static int popcount32(uint32_t v) { return __builtin_popcount(v); }

double or_bend_scale(double ia, double ic, uint32_t maskA, uint32_t maskC, int B) {
  uint32_t full = B >= 32 ? 0xffffffffu : ((1u << B) - 1u);
  int cont = popcount32(~(maskA | maskC) & full);
  if (cont == 0) return 0;
  int ua = popcount32(~maskA & full), uc = popcount32(~maskC & full);
  double ra = ua > 0 ? (ia * B) / ua : 0, rc = uc > 0 ? (ic * B) / uc : 0;
  return ((double)cont / B) * fmin(ra, rc);
}

static double shear_cos0(void) { return sqrt(1 - pow(1 - (1 - SHEAR_JAM) * (1 - SHEAR_JAM), 2)); }

double or_shear_jam(double area) {
  double c = shear_cos0() * sqrt(fmin(1, fmax(0.02, area)));
  return 1 - sqrt(1 - sqrt(1 - c * c));
}

double or_cell_area(double cu, double tu, double cv, double rv) { return fmax(0.5 * (cu + tu), 0.5 * (cv + rv)); }

/* shearLaw (laws.ts:96-115): out = [force, stiffness, energy] */
void or_shear_law(double ks, double Ld, double l, double jam, double out[3]) {
  double d = l - Ld;
  double f = ks * d, k = ks, e = 0.5 * ks * d * d;
  double c = -d / Ld - jam;
  if (c > 0) {
    double extra = ks * (SHEAR_LOCK - 1);
    if (c < SHEAR_RAMP) {
      f -= (extra * Ld * c * c) / (2 * SHEAR_RAMP);
      k += (extra * c) / SHEAR_RAMP;
      e += (extra * Ld * Ld * c * c * c) / (6 * SHEAR_RAMP);
    } else {
      f -= extra * Ld * (c - 0.5 * SHEAR_RAMP);
      k += extra;
      e += extra * Ld * Ld * ((SHEAR_RAMP * SHEAR_RAMP) / 6 + 0.5 * (c - SHEAR_RAMP) * (c - SHEAR_RAMP) + 0.5 * SHEAR_RAMP * (c - SHEAR_RAMP));
    }
  }
  out[0] = f;
  out[1] = k;
  out[2] = e;
}

/* ---------------------------------------------------------------- stencil operator */

typedef struct {
  int nx, ny, offset;
} lay_t;

typedef struct {
  int N, nl;
  lay_t lay[8];
  double* A;        /* N * 25 * 9 */
  uint8_t* fixed;   /* N */
  int* ij;          /* 3N */
  uint8_t* link;    /* fine level only: bit 0 +u, 1 -u, 2 +v, 3 -v connectivity (NULL: from A) */
} op_t;

typedef struct {
  int* idx;   /* 4N */
  double* w;  /* 4N */
} prol_t;

static void op_init(op_t* A, const lay_t* lay, int nl) {
  memset(A, 0, sizeof(*A));
  A->nl = nl;
  int N = 0;
  for (int l = 0; l < nl; l++) {
    A->lay[l] = lay[l];
    N += lay[l].nx * lay[l].ny;
  }
  A->N = N;
  A->A = calloc((size_t)N * STENCIL * 9, sizeof(double));
  A->fixed = calloc(N, 1);
  A->ij = malloc(sizeof(int) * 3 * N);
  for (int l = 0; l < nl; l++)
    for (int j = 0; j < lay[l].ny; j++)
      for (int i = 0; i < lay[l].nx; i++) {
        int n = lay[l].offset + j * lay[l].nx + i;
        A->ij[3 * n] = i;
        A->ij[3 * n + 1] = j;
        A->ij[3 * n + 2] = l;
      }
}

static void op_free(op_t* A) {
  free(A->A);
  free(A->fixed);
  free(A->ij);
  free(A->link);
  memset(A, 0, sizeof(*A));
}

static inline int op_nbr(const op_t* A, int n, int di, int dj) {
  int i = A->ij[3 * n] + di, j = A->ij[3 * n + 1] + dj;
  const lay_t* L = &A->lay[A->ij[3 * n + 2]];
  if (i < 0 || j < 0 || i >= L->nx || j >= L->ny) return -1;
  return L->offset + j * L->nx + i;
}

static inline void add_block(op_t* A, int row, int o, double s, const double* B) {
  double* d = A->A + ((size_t)row * STENCIL + o) * 9;
  for (int q = 0; q < 9; q++) d[q] += s * B[q];
}

/* StencilOperator.apply (newtonMultigrid.ts:95-124) */
static void op_apply(const op_t* A, const double* x, double* y) {
  /* rows are independent: threads change no result */
#pragma omp parallel for schedule(static)
  for (int n = 0; n < A->N; n++) {
    if (A->fixed[n]) {
      y[3 * n] = x[3 * n];
      y[3 * n + 1] = x[3 * n + 1];
      y[3 * n + 2] = x[3 * n + 2];
      continue;
    }
    double a0 = 0, a1 = 0, a2 = 0;
    for (int dj = -2; dj <= 2; dj++)
      for (int di = -2; di <= 2; di++) {
        int m = op_nbr(A, n, di, dj);
        if (m < 0 || A->fixed[m]) continue;
        const double* b = A->A + ((size_t)n * STENCIL + offset_index(di, dj)) * 9;
        double x0 = x[3 * m], x1 = x[3 * m + 1], x2 = x[3 * m + 2];
        a0 += b[0] * x0 + b[1] * x1 + b[2] * x2;
        a1 += b[3] * x0 + b[4] * x1 + b[5] * x2;
        a2 += b[6] * x0 + b[7] * x1 + b[8] * x2;
      }
    y[3 * n] = a0;
    y[3 * n + 1] = a1;
    y[3 * n + 2] = a2;
  }
}

/* ---------------------------------------------------------------- solver state */

typedef struct {
  int N, nl;
  bd_layer lay[8];
  bd_params P;
  /* scene statics */
  double* mass;
  uint8_t* pinned;
  double* cn0;
  float* fstrength;
  uint8_t* farea;
  int C;
  /* state */
  double *x, *v, *dx, *pred, *rest, *intact, *fail_site, *tails, *strain, *last_update, *softening;
  uint32_t* cut_mask;
  uint16_t* seg_flags;
  uint8_t* damage;
  int* nbr;
  uint8_t* layer_of;
  /* grips */
  int grip_count;
  int* grip_of;    /* per node: grip id or -1 */
  double* grip_w;  /* per node */
  double* grip_off;/* 3 per node */
  double goal[BD_MAX_GRIPS][3];
  double gk[BD_MAX_GRIPS];
  /* multigrid */
  int nlev;
  op_t lv[MAXLV];
  prol_t pr[MAXLV];
  int colors[MAXLV];  /* line colours per family of each level's smoother (0 = 3) */
  int lagged;
  /* counters */
  int broken, pulled, rejections;
  int pcg_iters;
  double pcg_relres;
  double last_alpha, last_e[4], last_m[4];
  /* external node terms and contact bounds, all off by default (assemble.wgsl:183-215,305-322;
     energy.wgsl:106-126; step.wgsl:61-112): ext 9 per node (gradient xyz, Hessian xx xy xz yy yz zz),
     cmeta 8 per node (own and reciprocal bound: direction xyz, largest approach), the floor plane */
  double* ext;
  double* cmeta;
  int floor_on;
  double floor_z;
  /* exact coarsest solve (BD_COARSE_EXACT): banded Cholesky factor of the coarsest level */
  int cn, cbw;
  double* cL;
  /* detached pieces' curved line-search steps (energy.wgsl:65-73), 7 per node: rotation w xyz,
     flag, arm r xyz; NULL when no node belongs to a piece */
  double* curve;
} or_solver;

static inline const bd_material* mat_of(const or_solver* S, int n) { return &S->lay[S->layer_of[n]].material; }

/* ---------------------------------------------------------------- external terms, floor, bounds */

#define FLOOR_BULGE 0.3e-3
#define FLOOR_K 1300.0
#define FLOOR_MU 0.35
#define FLOOR_SPEED 0.002
#define ACTIVE_BOUND 100.0

/* ext and floor terms of node n at the iterate: adds the gradient to g and the symmetric Hessian
   [xx xy xz yy yz zz] to H; returns whether the node sits on the floor's hard limit */
static int external_terms(const or_solver* S, int n, double* g, double* H) {
  if (S->ext) {
    for (int q = 0; q < 3; q++) g[q] += S->ext[9 * n + q];
    for (int q = 0; q < 6; q++) H[q] += S->ext[9 * n + 3 + q];
  }
  int core = 0;
  if (S->floor_on) {
    double half = 0.5 * mat_of(S, n)->thickness, delta = half + FLOOR_BULGE;
    double zc = S->x[3 * n + 2] + S->dx[3 * n + 2];
    double s = delta - (zc - S->floor_z);
    if (s > 0) {
      double lam = FLOOR_K * s * s / delta;
      double ux = S->dx[3 * n], uy = S->dx[3 * n + 1];
      double kf = FLOOR_MU * lam / fmax(sqrt(ux * ux + uy * uy), FLOOR_SPEED * S->P.h);
      g[0] += kf * ux;
      g[1] += kf * uy;
      g[2] -= lam;
      H[0] += kf;
      H[3] += kf;
      H[5] += 2 * FLOOR_K * s / delta;
      core = zc <= S->floor_z + half * (1 + 1e-3);
    }
  }
  return core;
}

/* bound / feasible (step.wgsl:61-77) in FP64 */
static void bound_d(double* d, const double* m) {
  double approach = -(d[0] * m[0] + d[1] * m[1] + d[2] * m[2]);
  if (approach > m[3])
    for (int q = 0; q < 3; q++) d[q] += (approach - m[3]) * m[q];
}

/* floorBound (step.wgsl:79-83 with common.wgsl floorLowest, floorWeightDepth) */
static double floor_bound_z(const or_solver* S, int n, double linz, double dz) {
  double half = 0.5 * mat_of(S, n)->thickness, delta = half + FLOOR_BULGE;
  double g = sqrt(S->P.gravity[0] * S->P.gravity[0] + S->P.gravity[1] * S->P.gravity[1] + S->P.gravity[2] * S->P.gravity[2]);
  double sW = sqrt(S->mass[n] * g * delta / FLOOR_K);
  double z0 = S->x[3 * n + 2], zl = z0 + linz;
  double s = fmax(delta - (zl - S->floor_z), 0);
  double lowest = fmin(fmax(S->floor_z + delta - fmax(sW, 2 * s), S->floor_z + half), zl);
  return fmax(dz, lowest - z0);
}

/* the step d of node n (from the iterate step lin) made feasible for its bounds and the floor */
static void feasible_d(const or_solver* S, int n, const double* lin, double* d) {
  if (S->cmeta) {
    bound_d(d, &S->cmeta[8 * n]);
    bound_d(d, &S->cmeta[8 * n + 4]);
    bound_d(d, &S->cmeta[8 * n]);
  }
  if (S->floor_on) d[2] = floor_bound_z(S, n, lin[2], d[2]);
}

static inline int seg_of(const or_solver* S, int n, int f, int dir) {
  if (dir > 0) return S->nbr[4 * n + 2 * f + 1] >= 0 ? 2 * n + f : -1;
  int p = S->nbr[4 * n + 2 * f];
  return p >= 0 ? 2 * p + f : -1;
}

static inline int seg_active(const or_solver* S, int s) {
  return s >= 0 && (S->seg_flags[s] & (SEG_EXISTS | SEG_BROKEN)) == SEG_EXISTS;
}

static inline int seg_other(const or_solver* S, int s) { return S->nbr[4 * (s >> 1) + 2 * (s & 1) + 1]; }

static int diag_active(const or_solver* S, int c, int d) {
  int r = S->nbr[4 * c + 1], t = S->nbr[4 * c + 3];
  if (r < 0 || t < 0) return 0;
  int cu = seg_active(S, 2 * c), cv = seg_active(S, 2 * c + 1);
  int rv = seg_active(S, 2 * r + 1), tu = seg_active(S, 2 * t);
  return d == 0 ? ((cu && rv) || (cv && tu)) : ((cu && cv) || (rv && tu));
}

static inline void posn(const or_solver* S, int n, double* o) {
  o[0] = S->x[3 * n] + S->dx[3 * n];
  o[1] = S->x[3 * n + 1] + S->dx[3 * n + 1];
  o[2] = S->x[3 * n + 2] + S->dx[3 * n + 2];
}

static inline double norm3(const double* a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

static void span_response(const or_solver* S, int s, double eps, double* T, double* K) {
  const bd_yarn* y = &mat_of(S, s >> 1)->yarns[s & 1];
  *T = S->intact[s] * or_axial_tension(y, eps);
  *K = S->intact[s] * or_axial_stiffness(y, eps);
}

static double seg_length(const or_solver* S, int s) {
  int a = s >> 1, b = seg_other(S, s);
  double d[3] = {S->x[3 * b] - S->x[3 * a], S->x[3 * b + 1] - S->x[3 * a + 1], S->x[3 * b + 2] - S->x[3 * a + 2]};
  return norm3(d);
}

/* airDrag (solver.ts:132-165) */
static void air_drag(const or_solver* S, int n, double D[9]) {
  memset(D, 0, 9 * sizeof(double));
  const bd_material* m = mat_of(S, n);
  if (!(m->air_drag > 0)) return;
  double k = (m->air_drag * m->pitch_u * m->pitch_v) / S->P.h;
  double t[2][3], l[2];
  for (int f = 0; f < 2; f++) {
    int a = n, b = n;
    if (seg_active(S, seg_of(S, n, f, -1))) a = S->nbr[4 * n + 2 * f];
    if (seg_active(S, seg_of(S, n, f, 1))) b = S->nbr[4 * n + 2 * f + 1];
    for (int q = 0; q < 3; q++) t[f][q] = S->x[3 * b + q] - S->x[3 * a + q];
    l[f] = norm3(t[f]);
  }
  double e[3], kPar, kPerp;
  if (l[0] > 0 && l[1] > 0) {
    double c[3] = {t[0][1] * t[1][2] - t[0][2] * t[1][1], t[0][2] * t[1][0] - t[0][0] * t[1][2], t[0][0] * t[1][1] - t[0][1] * t[1][0]};
    double lc = norm3(c);
    if (lc > 1e-3 * l[0] * l[1]) {
      for (int q = 0; q < 3; q++) e[q] = c[q] / lc;
      kPar = k;
      kPerp = 0;
      goto outer;
    }
  }
  if (l[0] > 0 || l[1] > 0) {
    int f = l[0] >= l[1] ? 0 : 1;
    for (int q = 0; q < 3; q++) e[q] = t[f][q] / l[f];
    kPar = 0;
    kPerp = k;
    goto outer;
  }
  e[0] = 1;
  e[1] = 0;
  e[2] = 0;
  kPar = k;
  kPerp = k;
outer:
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) D[3 * r + c] = (kPar - kPerp) * e[r] * e[c] + (r == c ? kPerp : 0);
}

static void drag_mul(const double* D, const double* v, double* o) {
  o[0] = D[0] * v[0] + D[1] * v[1] + D[2] * v[2];
  o[1] = D[3] * v[0] + D[4] * v[1] + D[5] * v[2];
  o[2] = D[6] * v[0] + D[7] * v[1] + D[8] * v[2];
}

/* gripTerm (solver.ts:173-180), generalised to several grips with disjoint patches */
static double grip_term(const or_solver* S, int n, double* g) {
  int k = S->grip_of[n];
  if (k < 0) return 0;
  double kg = S->gk[k] * S->grip_w[n];
  for (int q = 0; q < 3; q++) g[q] += kg * (S->x[3 * n + q] + S->dx[3 * n + q] - (S->goal[k][q] + S->grip_off[3 * n + q]));
  return kg;
}

static double grip_energy(const or_solver* S, int n) {
  int k = S->grip_of[n];
  if (k < 0) return 0;
  double l2 = 0;
  for (int q = 0; q < 3; q++) {
    double d = S->x[3 * n + q] + S->dx[3 * n + q] - (S->goal[k][q] + S->grip_off[3 * n + q]);
    l2 += d * d;
  }
  return 0.5 * S->gk[k] * S->grip_w[n] * l2;
}

/* segForce (solver.ts:254-277): returns fs, fills e (unit), kpar, kperp, l */
static double seg_force(const or_solver* S, int s, double* e, double* kpar, double* kperp, double* lout) {
  int a = s >> 1, f = s & 1, b = seg_other(S, s);
  const double* x = S->x;
  const double* dx = S->dx;
  double en[3], ev[3];
  for (int q = 0; q < 3; q++) {
    en[q] = x[3 * b + q] - x[3 * a + q];
    ev[q] = en[q] + (dx[3 * b + q] - dx[3 * a + q]);
  }
  double l = norm3(ev), ln = norm3(en);
  for (int q = 0; q < 3; q++) e[q] = ev[q] / l;
  double L = S->rest[s];
  const bd_yarn* y = &mat_of(S, a)->yarns[f];
  double T, K;
  span_response(S, s, l / L - 1, &T, &K);
  double c = y->axial_viscosity / (L * S->P.h);
  double fs = T + c * (l - ln);
  if (kpar) *kpar = fmax(K, 0) / L + c;
  if (kperp) *kperp = fmax(fs, 0) / l;
  if (lout) *lout = l;
  return fs;
}

/* bendCoeffs (solver.ts:284-291): kb, kd, fold, wa, wc */
static void bend_coeffs(const or_solver* S, int b, int f, int s0, int s1, double* kb, double* kd, double* fold, double* wa, double* wc) {
  const bd_yarn* y = &mat_of(S, b)->yarns[f];
  double g[4];
  or_bend_geometry(S->rest[s0], S->rest[s1], g);
  double sc = or_bend_scale(S->intact[s0], S->intact[s1], S->cut_mask[s0], S->cut_mask[s1], y->yarns_per_span);
  *kb = y->bend_rigidity * sc * g[2];
  *kd = (*kb * y->bend_viscosity) / S->P.h;
  *fold = g[3];
  *wa = g[0];
  *wc = g[1];
}

static void triple(const or_solver* S, int a, int b, int c, double wa, double wc, double* r, double* rn) {
  for (int k = 0; k < 3; k++) {
    rn[k] = wa * S->x[3 * a + k] - 2 * S->x[3 * b + k] + wc * S->x[3 * c + k];
    r[k] = rn[k] + wa * S->dx[3 * a + k] - 2 * S->dx[3 * b + k] + wc * S->dx[3 * c + k];
  }
}

static double bend_grad_coeff(double kb, double fold, const double* r) {
  double ps[3];
  or_bend_psi(dot3(r, r), fold, ps);
  return kb * ps[1];
}

static double bend_visc_coeff(double kd, const double* r, const double* rn) {
  double lr = norm3(r);
  return (kd * (lr - norm3(rn))) / fmax(lr, 1e-12);
}

/* addBendingGradient (solver.ts:319-352) */
static void add_bending_gradient(const or_solver* S, int n, int f, double* g) {
  int prev = S->nbr[4 * n + 2 * f], nxt = S->nbr[4 * n + 2 * f + 1];
  int sPrev = prev >= 0 ? 2 * prev + f : -1;
  int sNext = nxt >= 0 ? 2 * n + f : -1;
  double r[3], rn[3], kb, kd, ab, wa, wc;
  if (seg_active(S, sPrev) && seg_active(S, sNext)) {
    bend_coeffs(S, n, f, sPrev, sNext, &kb, &kd, &ab, &wa, &wc);
    triple(S, prev, n, nxt, wa, wc, r, rn);
    double kv = bend_grad_coeff(kb, ab, r) + bend_visc_coeff(kd, r, rn);
    for (int k = 0; k < 3; k++) g[k] += -2 * kv * r[k];
  }
  if (seg_active(S, sNext)) {
    int nxt2 = S->nbr[4 * nxt + 2 * f + 1];
    int sNext2 = nxt2 >= 0 ? 2 * nxt + f : -1;
    if (seg_active(S, sNext2)) {
      bend_coeffs(S, nxt, f, sNext, sNext2, &kb, &kd, &ab, &wa, &wc);
      triple(S, n, nxt, nxt2, wa, wc, r, rn);
      double kv = bend_grad_coeff(kb, ab, r) + bend_visc_coeff(kd, r, rn);
      for (int k = 0; k < 3; k++) g[k] += wa * kv * r[k];
    }
  }
  if (seg_active(S, sPrev)) {
    int prev2 = S->nbr[4 * prev + 2 * f];
    int sPrev2 = prev2 >= 0 ? 2 * prev2 + f : -1;
    if (seg_active(S, sPrev2)) {
      bend_coeffs(S, prev, f, sPrev2, sPrev, &kb, &kd, &ab, &wa, &wc);
      triple(S, prev2, prev, n, wa, wc, r, rn);
      double kv = bend_grad_coeff(kb, ab, r) + bend_visc_coeff(kd, r, rn);
      for (int k = 0; k < 3; k++) g[k] += wc * kv * r[k];
    }
  }
}

/* cellShear (solver.ts:374-380) */
static void cell_shear(const or_solver* S, int cell, double* ks, double* jam) {
  int r = S->nbr[4 * cell + 1], t = S->nbr[4 * cell + 3];
#define IO(s) (seg_active(S, (s)) ? S->intact[(s)] : 0.0)
  const bd_material* m = mat_of(S, cell);
  *ks = (m->shear_stiffness * S->cn0[cell]) / m->crossing_normal0;
  *jam = or_shear_jam(or_cell_area(IO(2 * cell), IO(2 * t), IO(2 * cell + 1), IO(2 * r + 1)));
#undef IO
}

static double shear_rest(const or_solver* S, int cell) { return hypot(S->rest[2 * cell], S->rest[2 * cell + 1]); }

/* nodeGradient (solver.ts:387-422) */
void or_node_gradient(or_solver* S, int n, double* g) {
  g[0] = g[1] = g[2] = 0;
  if (S->pinned[n]) return;
  double h = S->P.h;
  double w = S->mass[n] / (h * h);
  for (int q = 0; q < 3; q++) g[q] += w * (S->dx[3 * n + q] - S->pred[3 * n + q]);
  grip_term(S, n, g);
  double Hx[6] = {0, 0, 0, 0, 0, 0};
  external_terms(S, n, g, Hx);
  double D[9], Dg[3];
  air_drag(S, n, D);
  drag_mul(D, &S->dx[3 * n], Dg);
  for (int q = 0; q < 3; q++) g[q] += Dg[q];
  double e[3];
  for (int f = 0; f < 2; f++) {
    int sp = seg_of(S, n, f, -1), sn = seg_of(S, n, f, 1);
    if (seg_active(S, sp)) {
      double fs = seg_force(S, sp, e, 0, 0, 0);
      for (int q = 0; q < 3; q++) g[q] += fs * e[q];
    }
    if (seg_active(S, sn)) {
      double fs = seg_force(S, sn, e, 0, 0, 0);
      for (int q = 0; q < 3; q++) g[q] -= fs * e[q];
    }
    add_bending_gradient(S, n, f, g);
  }
  /* shearPartners (solver.ts:355-367) */
  const int* nb = S->nbr;
  int left = nb[4 * n], right = nb[4 * n + 1], down = nb[4 * n + 2], up = nb[4 * n + 3];
  int pc[4], pp[4], pd[4], np = 0;
  if (right >= 0 && up >= 0) { pc[np] = n; pp[np] = nb[4 * up + 1]; pd[np++] = 0; }
  if (left >= 0 && down >= 0) { pc[np] = nb[4 * down]; pp[np] = nb[4 * down]; pd[np++] = 0; }
  if (left >= 0 && up >= 0) { pc[np] = left; pp[np] = nb[4 * up]; pd[np++] = 1; }
  if (right >= 0 && down >= 0) { pc[np] = down; pp[np] = nb[4 * down + 1]; pd[np++] = 1; }
  for (int k = 0; k < np; k++) {
    if (pp[k] < 0 || !diag_active(S, pc[k], pd[k])) continue;
    double ks, jam, pa[3], pb[3], sl[3];
    cell_shear(S, pc[k], &ks, &jam);
    posn(S, n, pa);
    posn(S, pp[k], pb);
    double d0 = pb[0] - pa[0], d1 = pb[1] - pa[1], d2 = pb[2] - pa[2];
    double l = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    or_shear_law(ks, shear_rest(S, pc[k]), l, jam, sl);
    g[0] -= (sl[0] * d0) / l;
    g[1] -= (sl[0] * d1) / l;
    g[2] -= (sl[0] * d2) / l;
  }
}

/* lenChange (solver.ts:447-451) */
static void len_change(const double* e0, const double* de, double a, double l0, double* dl, double* la) {
  double ea[3] = {e0[0] + a * de[0], e0[1] + a * de[1], e0[2] + a * de[2]};
  *la = norm3(ea);
  *dl = (2 * a * dot3(e0, de) + a * a * dot3(de, de)) / fmax(*la + l0, 1e-12);
}

/* rotationCurve (common.wgsl:277-283): what a rotation by th adds to its tangent th x r */
static void rotation_curve(const double* th, const double* r, double* o) {
  double phi = norm3(th);
  if (phi < 1e-9) {
    o[0] = o[1] = o[2] = 0;
    return;
  }
  double k[3] = {th[0] / phi, th[1] / phi, th[2] / phi}, kr[3], kkr[3];
  cross3(k, r, kr);
  cross3(k, kr, kkr);
  double a = sin(phi) - phi, c = 1 - cos(phi);
  for (int q = 0; q < 3; q++) o[q] = a * kr[q] + c * kkr[q];
}

static void energy_change_curved(or_solver* S, const double* d, double* out, double* mag);

/* energyChange (solver.ts:430-534) for the 4 line-search alphas */
void or_energy_change(or_solver* S, const double* d, double* out, double* mag) {
  if (S->curve) {
    energy_change_curved(S, d, out, mag);
    return;
  }
  const int NA = 4;
  const double* al = LINE_ALPHAS;
  for (int k = 0; k < NA; k++) out[k] = mag[k] = 0;
  double h = S->P.h;
  const double* x = S->x;
  const double* dx = S->dx;
  static const double Z[3] = {0, 0, 0};
#define DIR(n) (S->pinned[n] ? Z : &d[3 * (n)])
#define ADD(k, v) do { double v_ = (v); out[k] += v_; mag[k] += fabs(v_); } while (0)
#define QUAD(gd, qq) do { double gd_ = (gd), q_ = (qq); for (int k = 0; k < NA; k++) { out[k] += al[k] * gd_ + 0.5 * al[k] * al[k] * q_; mag[k] += fabs(al[k] * gd_) + fabs(0.5 * al[k] * al[k] * q_); } } while (0)
  for (int n = 0; n < S->N; n++) {
    const double* dn = DIR(n);
    const double* d0 = &dx[3 * n];
    if (!S->pinned[n]) {
      double w = S->mass[n] / (h * h);
      double t = dn[0] * (d0[0] - S->pred[3 * n]) + dn[1] * (d0[1] - S->pred[3 * n + 1]) + dn[2] * (d0[2] - S->pred[3 * n + 2]);
      QUAD(w * t, w * dot3(dn, dn));
      double g[3] = {0, 0, 0};
      double kg = grip_term(S, n, g);
      if (kg > 0) QUAD(g[0] * dn[0] + g[1] * dn[1] + g[2] * dn[2], kg * dot3(dn, dn));
      double D[9], t1[3], t2[3];
      air_drag(S, n, D);
      drag_mul(D, d0, t1);
      drag_mul(D, dn, t2);
      QUAD(dot3(t1, dn), dot3(dn, t2));
      if (S->ext || S->floor_on) {
        /* external terms and the floor: quadratic models at this iterate (energy.wgsl:106-126) */
        double ge[3] = {0, 0, 0}, h6[6] = {0, 0, 0, 0, 0, 0};
        external_terms(S, n, ge, h6);
        double Hd[3] = {h6[0] * dn[0] + h6[1] * dn[1] + h6[2] * dn[2], h6[1] * dn[0] + h6[3] * dn[1] + h6[4] * dn[2],
                        h6[2] * dn[0] + h6[4] * dn[1] + h6[5] * dn[2]};
        QUAD(dot3(ge, dn), dot3(dn, Hd));
      }
    }
    const bd_material* mat = mat_of(S, n);
    for (int f = 0; f < 2; f++) {
      const bd_yarn* yarn = &mat->yarns[f];
      int next = S->nbr[4 * n + 2 * f + 1];
      int s = 2 * n + f;
      if (next >= 0 && seg_active(S, s)) {
        double L = S->rest[s];
        double en[3], ec[3], de[3];
        const double* db = DIR(next);
        for (int q = 0; q < 3; q++) {
          en[q] = x[3 * next + q] - x[3 * n + q];
          ec[q] = en[q] + dx[3 * next + q] - d0[q];
          de[q] = db[q] - dn[q];
        }
        double l0 = norm3(ec), ln = norm3(en);
        double T0, K0;
        span_response(S, s, l0 / L - 1, &T0, &K0);
        double c = yarn->axial_viscosity / (L * h);
        for (int k = 0; k < NA; k++) {
          double dl, la, Ta, Ka;
          len_change(ec, de, al[k], l0, &dl, &la);
          span_response(S, s, la / L - 1, &Ta, &Ka);
          ADD(k, 0.5 * (T0 + Ta) * dl);
          ADD(k, 0.5 * c * dl * (la + l0 - 2 * ln));
        }
      }
      int prev = S->nbr[4 * n + 2 * f];
      if (prev >= 0 && next >= 0 && seg_active(S, 2 * prev + f) && seg_active(S, s)) {
        double kb, kd, ab, wa, wc;
        bend_coeffs(S, n, f, 2 * prev + f, s, &kb, &kd, &ab, &wa, &wc);
        double rn[3], r0[3], dr[3];
        const double* dp = DIR(prev);
        const double* dq = DIR(next);
        for (int q = 0; q < 3; q++) {
          rn[q] = wa * x[3 * prev + q] - 2 * x[3 * n + q] + wc * x[3 * next + q];
          r0[q] = rn[q] + wa * dx[3 * prev + q] - 2 * d0[q] + wc * dx[3 * next + q];
          dr[q] = wa * dp[q] - 2 * dn[q] + wc * dq[q];
        }
        double q0 = dot3(r0, r0), l0 = sqrt(q0), lrn = norm3(rn);
        for (int k = 0; k < NA; k++) {
          double a = al[k];
          double dqa = 2 * a * dot3(r0, dr) + a * a * dot3(dr, dr);
          ADD(k, 0.5 * kb * or_bend_psi_delta(q0, q0 + dqa, dqa, ab));
          double dl, la;
          len_change(r0, dr, a, l0, &dl, &la);
          ADD(k, 0.5 * kd * dl * (la + l0 - 2 * lrn));
        }
      }
    }
    if (S->nbr[4 * n + 1] >= 0 && S->nbr[4 * n + 3] >= 0) {
      int r = S->nbr[4 * n + 1], t = S->nbr[4 * n + 3], rt = S->nbr[4 * r + 3];
      double Ld = shear_rest(S, n), ks, jam;
      cell_shear(S, n, &ks, &jam);
      int P2[2][2] = {{n, rt}, {r, t}};
      for (int dgi = 0; dgi < 2; dgi++) {
        if (!diag_active(S, n, dgi)) continue;
        int p = P2[dgi][0], q2 = P2[dgi][1];
        double ec[3], de[3];
        const double* dq = DIR(q2);
        const double* dp = DIR(p);
        for (int z = 0; z < 3; z++) {
          ec[z] = x[3 * q2 + z] + dx[3 * q2 + z] - x[3 * p + z] - dx[3 * p + z];
          de[z] = dq[z] - dp[z];
        }
        double l0 = norm3(ec), sl[3], sa[3];
        or_shear_law(ks, Ld, l0, jam, sl);
        for (int k = 0; k < NA; k++) {
          double dl, la;
          len_change(ec, de, al[k], l0, &dl, &la);
          or_shear_law(ks, Ld, la, jam, sa);
          ADD(k, 0.5 * (sl[0] + sa[0]) * dl);
        }
      }
    }
  }
#undef DIR
#undef ADD
#undef QUAD
}

/* The same with detached pieces' nodes stepping along their piece's rotation (energy.wgsl:65-73):
   at step length a a node's step is a d + rotationCurve(a w, r) where its curve flag is set. Every
   term is evaluated from the nodes' steps: node-local quadratic models as g . s + s^T H s / 2, span,
   bending and shear lengths from the displacement of their vectors. */
static void len_change_d(const double* e0, const double* D, double l0, double* dl, double* la) {
  double ea[3] = {e0[0] + D[0], e0[1] + D[1], e0[2] + D[2]};
  *la = norm3(ea);
  *dl = (2 * dot3(e0, D) + dot3(D, D)) / fmax(*la + l0, 1e-12);
}

static void energy_change_curved(or_solver* S, const double* d, double* out, double* mag) {
  const int NA = 4, N = S->N;
  const double* al = LINE_ALPHAS;
  for (int k = 0; k < NA; k++) out[k] = mag[k] = 0;
  const double h = S->P.h;
  const double* x = S->x;
  const double* dx = S->dx;
  double* st = calloc((size_t)N * NA * 3, sizeof(double));
  for (int n = 0; n < N; n++) {
    if (S->pinned[n]) continue;
    const double* cv = &S->curve[7 * (size_t)n];
    for (int k = 0; k < NA; k++) {
      double* o = &st[((size_t)n * NA + k) * 3];
      for (int q = 0; q < 3; q++) o[q] = al[k] * d[3 * n + q];
      if (cv[3] >= 0.5) {
        double th[3] = {al[k] * cv[0], al[k] * cv[1], al[k] * cv[2]}, c[3];
        rotation_curve(th, &cv[4], c);
        for (int q = 0; q < 3; q++) o[q] += c[q];
      }
    }
  }
#define S_(n, k) (&st[((size_t)(n) * NA + (k)) * 3])
#define ADD(k, v) do { double v_ = (v); out[k] += v_; mag[k] += fabs(v_); } while (0)
  /* g . s + s^T H s / 2 for a symmetric H given as 9 entries */
#define QUADS(n, g, H9) do { for (int k = 0; k < NA; k++) { const double* s_ = S_(n, k); double Hs_[3]; \
    drag_mul(H9, s_, Hs_); ADD(k, dot3(g, s_)); ADD(k, 0.5 * dot3(s_, Hs_)); } } while (0)
  for (int n = 0; n < N; n++) {
    const double* d0 = &dx[3 * n];
    if (!S->pinned[n]) {
      double w = S->mass[n] / (h * h);
      double g[3] = {w * (d0[0] - S->pred[3 * n]), w * (d0[1] - S->pred[3 * n + 1]), w * (d0[2] - S->pred[3 * n + 2])};
      double Hw[9] = {w, 0, 0, 0, w, 0, 0, 0, w};
      QUADS(n, g, Hw);
      double gg[3] = {0, 0, 0};
      double kg = grip_term(S, n, gg);
      if (kg > 0) {
        double Hg[9] = {kg, 0, 0, 0, kg, 0, 0, 0, kg};
        QUADS(n, gg, Hg);
      }
      double D[9], t1[3];
      air_drag(S, n, D);
      drag_mul(D, d0, t1);
      QUADS(n, t1, D);
      if (S->ext || S->floor_on) {
        double ge[3] = {0, 0, 0}, h6[6] = {0, 0, 0, 0, 0, 0};
        external_terms(S, n, ge, h6);
        double He[9] = {h6[0], h6[1], h6[2], h6[1], h6[3], h6[4], h6[2], h6[4], h6[5]};
        QUADS(n, ge, He);
      }
    }
    const bd_material* mat = mat_of(S, n);
    for (int f = 0; f < 2; f++) {
      const bd_yarn* yarn = &mat->yarns[f];
      int next = S->nbr[4 * n + 2 * f + 1];
      int s = 2 * n + f;
      if (next >= 0 && seg_active(S, s)) {
        double L = S->rest[s];
        double en[3], ec[3];
        for (int q = 0; q < 3; q++) {
          en[q] = x[3 * next + q] - x[3 * n + q];
          ec[q] = en[q] + dx[3 * next + q] - d0[q];
        }
        double l0 = norm3(ec), ln = norm3(en);
        double T0, K0;
        span_response(S, s, l0 / L - 1, &T0, &K0);
        double c = yarn->axial_viscosity / (L * h);
        for (int k = 0; k < NA; k++) {
          double Dk[3], dl, la, Ta, Ka;
          for (int q = 0; q < 3; q++) Dk[q] = S_(next, k)[q] - S_(n, k)[q];
          len_change_d(ec, Dk, l0, &dl, &la);
          span_response(S, s, la / L - 1, &Ta, &Ka);
          ADD(k, 0.5 * (T0 + Ta) * dl);
          ADD(k, 0.5 * c * dl * (la + l0 - 2 * ln));
        }
      }
      int prev = S->nbr[4 * n + 2 * f];
      if (prev >= 0 && next >= 0 && seg_active(S, 2 * prev + f) && seg_active(S, s)) {
        double kb, kd, ab, wa, wc;
        bend_coeffs(S, n, f, 2 * prev + f, s, &kb, &kd, &ab, &wa, &wc);
        double rn[3], r0[3];
        for (int q = 0; q < 3; q++) {
          rn[q] = wa * x[3 * prev + q] - 2 * x[3 * n + q] + wc * x[3 * next + q];
          r0[q] = rn[q] + wa * dx[3 * prev + q] - 2 * d0[q] + wc * dx[3 * next + q];
        }
        double q0 = dot3(r0, r0), l0 = sqrt(q0), lrn = norm3(rn);
        for (int k = 0; k < NA; k++) {
          double DR[3], dl, la;
          for (int q = 0; q < 3; q++) DR[q] = wa * S_(prev, k)[q] - 2 * S_(n, k)[q] + wc * S_(next, k)[q];
          double dqa = 2 * dot3(r0, DR) + dot3(DR, DR);
          ADD(k, 0.5 * kb * or_bend_psi_delta(q0, q0 + dqa, dqa, ab));
          len_change_d(r0, DR, l0, &dl, &la);
          ADD(k, 0.5 * kd * dl * (la + l0 - 2 * lrn));
        }
      }
    }
    if (S->nbr[4 * n + 1] >= 0 && S->nbr[4 * n + 3] >= 0) {
      int r = S->nbr[4 * n + 1], t = S->nbr[4 * n + 3], rt = S->nbr[4 * r + 3];
      double Ld = shear_rest(S, n), ks, jam;
      cell_shear(S, n, &ks, &jam);
      int P2[2][2] = {{n, rt}, {r, t}};
      for (int dgi = 0; dgi < 2; dgi++) {
        if (!diag_active(S, n, dgi)) continue;
        int p = P2[dgi][0], q2 = P2[dgi][1];
        double ec[3];
        for (int z = 0; z < 3; z++) ec[z] = x[3 * q2 + z] + dx[3 * q2 + z] - x[3 * p + z] - dx[3 * p + z];
        double l0 = norm3(ec), sl[3], sa[3];
        or_shear_law(ks, Ld, l0, jam, sl);
        for (int k = 0; k < NA; k++) {
          double Dk[3], dl, la;
          for (int z = 0; z < 3; z++) Dk[z] = S_(q2, k)[z] - S_(p, k)[z];
          len_change_d(ec, Dk, l0, &dl, &la);
          or_shear_law(ks, Ld, la, jam, sa);
          ADD(k, 0.5 * (sl[0] + sa[0]) * dl);
        }
      }
    }
  }
#undef S_
#undef ADD
#undef QUADS
  free(st);
}

void or_set_curve(or_solver* S, const double* curve) {
  free(S->curve);
  S->curve = NULL;
  if (!curve) return;
  S->curve = malloc(sizeof(double) * 7 * (size_t)S->N);
  memcpy(S->curve, curve, sizeof(double) * 7 * (size_t)S->N);
}

/* potential (solver.ts:537-595) */
double or_potential(or_solver* S) {
  double h = S->P.h, E = 0;
  for (int n = 0; n < S->N; n++) {
    if (S->pinned[n]) continue;
    double w = S->mass[n] / (2 * h * h);
    for (int q = 0; q < 3; q++) {
      double d = S->dx[3 * n + q] - S->pred[3 * n + q];
      E += w * d * d;
    }
    E += grip_energy(S, n);
    double D[9], Dd[3];
    air_drag(S, n, D);
    drag_mul(D, &S->dx[3 * n], Dd);
    E += 0.5 * dot3(&S->dx[3 * n], Dd);
  }
  for (int s = 0; s < 2 * S->N; s++) {
    if (!seg_active(S, s)) continue;
    int a = s >> 1, f = s & 1, b = seg_other(S, s);
    double pa[3], pb[3], dd[3];
    posn(S, a, pa);
    posn(S, b, pb);
    for (int q = 0; q < 3; q++) dd[q] = pb[q] - pa[q];
    double l = norm3(dd), ln = seg_length(S, s), L = S->rest[s];
    const bd_yarn* y = &mat_of(S, a)->yarns[f];
    /* the reference's verification potential assumes history-free spans (intact = 1) */
    E += L * or_axial_energy_density(y, l / L - 1);
    double c = y->axial_viscosity / (L * h);
    E += 0.5 * c * (l - ln) * (l - ln);
    int nb = S->nbr[4 * b + 2 * f + 1];
    if (nb >= 0 && seg_active(S, 2 * b + f)) {
      double kb, kd, ab, wa, wc, r[3], rn[3], ps[3];
      bend_coeffs(S, b, f, s, 2 * b + f, &kb, &kd, &ab, &wa, &wc);
      triple(S, a, b, nb, wa, wc, r, rn);
      or_bend_psi(dot3(r, r), ab, ps);
      E += 0.5 * kb * ps[0];
      double dv = norm3(r) - norm3(rn);
      E += 0.5 * kd * dv * dv;
    }
  }
  for (int n = 0; n < S->N; n++) {
    if (S->nbr[4 * n + 1] < 0 || S->nbr[4 * n + 3] < 0) continue;
    int r = S->nbr[4 * n + 1], t = S->nbr[4 * n + 3], rt = S->nbr[4 * r + 3];
    double ks, jam, Ld = shear_rest(S, n);
    cell_shear(S, n, &ks, &jam);
    int P2[2][2] = {{n, rt}, {r, t}};
    for (int d = 0; d < 2; d++) {
      if (!diag_active(S, n, d)) continue;
      double pa[3], pb[3], dd[3], sl[3];
      posn(S, P2[d][0], pa);
      posn(S, P2[d][1], pb);
      for (int q = 0; q < 3; q++) dd[q] = pb[q] - pa[q];
      or_shear_law(ks, Ld, norm3(dd), jam, sl);
      E += sl[2];
    }
  }
  return E;
}

/* ---------------------------------------------------------------- assembly (newtonMultigrid.ts:169-274) */

static void axial_block(double kPar, double kPerp, const double* e, double* out) {
  double d = kPar - kPerp;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) out[r * 3 + c] = d * e[r] * e[c] + (r == c ? kPerp : 0);
}

static void scatter_pair(op_t* A, int a, int b, const double* H) {
  int ia = A->ij[3 * a], ja = A->ij[3 * a + 1], ib = A->ij[3 * b], jb = A->ij[3 * b + 1];
  if (!A->fixed[a]) {
    add_block(A, a, CENTER, 1, H);
    if (!A->fixed[b]) add_block(A, a, offset_index(ib - ia, jb - ja), -1, H);
  }
  if (!A->fixed[b]) {
    add_block(A, b, CENTER, 1, H);
    if (!A->fixed[a]) add_block(A, b, offset_index(ia - ib, ja - jb), -1, H);
  }
}

static const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};

static void solver_layers(const or_solver* S, lay_t* lay) {
  for (int l = 0; l < S->nl; l++) {
    lay[l].nx = S->lay[l].nx;
    lay[l].ny = S->lay[l].ny;
    lay[l].offset = S->lay[l].node_offset;
  }
}

static void fine_links(const or_solver* S, op_t* A) {
  A->link = calloc(S->N, 1);
  for (int n = 0; n < S->N; n++) {
    uint8_t m = 0;
    int s;
    s = seg_of(S, n, 0, 1);
    if (seg_active(S, s) && S->intact[s] > LINK_MIN_INTACT) m |= 1;
    s = seg_of(S, n, 0, -1);
    if (seg_active(S, s) && S->intact[s] > LINK_MIN_INTACT) m |= 2;
    s = seg_of(S, n, 1, 1);
    if (seg_active(S, s) && S->intact[s] > LINK_MIN_INTACT) m |= 4;
    s = seg_of(S, n, 1, -1);
    if (seg_active(S, s) && S->intact[s] > LINK_MIN_INTACT) m |= 8;
    A->link[n] = m;
  }
}

static void assemble_fine(or_solver* S, op_t* A) {
  lay_t lay[8];
  solver_layers(S, lay);
  op_init(A, lay, S->nl);
  fine_links(S, A);
  double h = S->P.h;
  for (int n = 0; n < S->N; n++)
    if (S->pinned[n]) {
      A->fixed[n] = 1;
      add_block(A, n, CENTER, 1, I3);
    }
  for (int n = 0; n < S->N; n++) {
    if (A->fixed[n]) continue;
    double g[3] = {0, 0, 0}, D[9];
    add_block(A, n, CENTER, S->mass[n] / (h * h) + grip_term(S, n, g), I3);
    air_drag(S, n, D);
    add_block(A, n, CENTER, 1, D);
  }
  double H[9], e[3];
  for (int s = 0; s < 2 * S->N; s++) {
    if (!seg_active(S, s)) continue;
    int a = s >> 1, b = seg_other(S, s);
    double kpar, kperp, l;
    seg_force(S, s, e, &kpar, &kperp, &l);
    axial_block(kpar, kperp, e, H);
    scatter_pair(A, a, b, H);
  }
  for (int b = 0; b < S->N; b++)
    for (int f = 0; f < 2; f++) {
      int a = S->nbr[4 * b + 2 * f], c = S->nbr[4 * b + 2 * f + 1];
      if (a < 0 || c < 0) continue;
      int s0 = 2 * a + f, s1 = 2 * b + f;
      if (!seg_active(S, s0) || !seg_active(S, s1)) continue;
      const bd_yarn* yarn = &mat_of(S, b)->yarns[f];
      double kb, kd, fold, wa, wc;
      bend_coeffs(S, b, f, s0, s1, &kb, &kd, &fold, &wa, &wc);
      double q2 = 0;
      for (int z = 0; z < 3; z++) {
        double rz = wa * (S->x[3 * a + z] + S->dx[3 * a + z]) - 2 * (S->x[3 * b + z] + S->dx[3 * b + z]) + wc * (S->x[3 * c + z] + S->dx[3 * c + z]);
        q2 += rz * rz;
      }
      double ps[3];
      or_bend_psi(q2, fold, ps);
      double k = kb * (ps[1] + 2 * ps[2] * q2) + (kb * yarn->bend_viscosity) / h;
      int nodes[3] = {a, b, c};
      double w[3] = {wa, -2, wc};
      for (int p = 0; p < 3; p++) {
        int r = nodes[p];
        if (A->fixed[r]) continue;
        for (int q = 0; q < 3; q++) {
          int m = nodes[q];
          if (A->fixed[m]) continue;
          int o = offset_index(A->ij[3 * m] - A->ij[3 * r], A->ij[3 * m + 1] - A->ij[3 * r + 1]);
          add_block(A, r, o, k * w[p] * w[q], I3);
        }
      }
    }
  for (int n = 0; n < S->N; n++) {
    int r = S->nbr[4 * n + 1], t = S->nbr[4 * n + 3];
    if (r < 0 || t < 0) continue;
    int rt = S->nbr[4 * r + 3];
    double ks, jam;
    cell_shear(S, n, &ks, &jam);
    double Ld = hypot(S->rest[2 * n], S->rest[2 * n + 1]);
    int P2[2][2] = {{n, rt}, {r, t}};
    for (int d = 0; d < 2; d++) {
      if (!diag_active(S, n, d)) continue;
      int p = P2[d][0], q = P2[d][1];
      double pa[3], l = 0, sl[3];
      for (int z = 0; z < 3; z++) {
        pa[z] = S->x[3 * q + z] + S->dx[3 * q + z] - S->x[3 * p + z] - S->dx[3 * p + z];
        l += pa[z] * pa[z];
      }
      l = sqrt(l);
      for (int z = 0; z < 3; z++) pa[z] /= l;
      or_shear_law(ks, Ld, l, jam, sl);
      axial_block(sl[1], fmax(sl[0], 0) / l, pa, H);
      scatter_pair(A, p, q, H);
    }
  }
  if (!S->ext && !S->cmeta && !S->floor_on) return;
  /* external node terms and the floor into the node's own block; then active bounds: a node at
     its bound whose net force still drives it on is held along the bound normal with ACTIVE_BOUND
     times the trace of its own block (assemble.wgsl:183-215, 305-322) */
  uint8_t* core = calloc((size_t)(S->N > 0 ? S->N : 1), 1);
  for (int n = 0; n < S->N; n++) {
    if (A->fixed[n]) continue;
    double g[3] = {0, 0, 0}, h6[6] = {0, 0, 0, 0, 0, 0};
    core[n] = (uint8_t)external_terms(S, n, g, h6);
    double B[9] = {h6[0], h6[1], h6[2], h6[1], h6[3], h6[4], h6[2], h6[4], h6[5]};
    add_block(A, n, CENTER, 1, B);
  }
  for (int n = 0; n < S->N; n++) {
    if (A->fixed[n]) continue;
    const double* D = A->A + ((size_t)n * STENCIL + CENTER) * 9;
    double kAct = ACTIVE_BOUND * (D[0] + D[4] + D[8]);
    double g[3];
    or_node_gradient(S, n, g);
    for (int b = 0; b < 2 && S->cmeta; b++) {
      const double* m = &S->cmeta[8 * n + 4 * b];
      if (m[3] >= 1e8 || (m[0] == 0 && m[1] == 0 && m[2] == 0)) continue;
      double approach = -dot3(&S->dx[3 * n], m);
      if (approach >= m[3] - 1e-3 * fabs(m[3]) && dot3(g, m) > 0) {
        axial_block(kAct, 0, m, H);
        add_block(A, n, CENTER, 1, H);
      }
    }
    if (core[n] && g[2] > 0) {
      static const double ez[3] = {0, 0, 1};
      axial_block(kAct, 0, ez, H);
      add_block(A, n, CENTER, 1, H);
    }
  }
  free(core);
}

/* ---------------------------------------------------------------- hierarchy */

static int coarse_count(int n) { return (n - 1 + 1) / 2 + 1; } /* ceil((n-1)/2) + 1 */
static int coarse_to_fine(int I, int n) { return 2 * I < n - 1 ? 2 * I : n - 1; }

/* bracket (lattice.ts:25-36): returns count, fills idx/w */
static int bracket(int i, int n, int* idx, double* w) {
  int nc = coarse_count(n);
  int lo = nc - 1 < (i >> 1) ? nc - 1 : (i >> 1);
  if (coarse_to_fine(lo, n) > i) lo--;
  int a = coarse_to_fine(lo, n);
  if (a == i) {
    idx[0] = lo;
    w[0] = 1;
    return 1;
  }
  int b = coarse_to_fine(lo + 1, n);
  idx[0] = lo;
  w[0] = (double)(b - i) / (b - a);
  idx[1] = lo + 1;
  w[1] = (double)(i - a) / (b - a);
  return 2;
}

static int should_coarsen(const op_t* A) {
  for (int l = 0; l < A->nl; l++)
    if (!(A->lay[l].nx * A->lay[l].ny > 48 && A->lay[l].nx >= 5 && A->lay[l].ny >= 5)) return 0;
  return 1;
}

static int linked(const op_t* A, int a, int di, int dj) {
  int b = op_nbr(A, a, di, dj);
  if (b < 0) return 0;
  if (A->link) {
    int bit = di > 0 ? 1 : di < 0 ? 2 : dj > 0 ? 4 : 8;
    return (A->link[a] & bit) != 0;
  }
  if (A->fixed[b]) return 1;
  const double* base = A->A + ((size_t)a * STENCIL + offset_index(di, dj)) * 9;
  for (int q = 0; q < 9; q++)
    if (base[q] != 0) return 1;
  return 0;
}

/* buildProlongation (newtonMultigrid.ts:305-342) */
static void build_prolongation(const op_t* F, lay_t* coarse, prol_t* P) {
  int off = 0;
  for (int l = 0; l < F->nl; l++) {
    coarse[l].nx = coarse_count(F->lay[l].nx);
    coarse[l].ny = coarse_count(F->lay[l].ny);
    coarse[l].offset = off;
    off += coarse[l].nx * coarse[l].ny;
  }
  P->idx = malloc(sizeof(int) * 4 * F->N);
  P->w = calloc(4 * F->N, sizeof(double));
  for (int q = 0; q < 4 * F->N; q++) P->idx[q] = -1;
  for (int l = 0; l < F->nl; l++) {
    const lay_t* L = &F->lay[l];
    const lay_t* C = &coarse[l];
    for (int j = 0; j < L->ny; j++) {
      int bj[2];
      double wj[2];
      int nj = bracket(j, L->ny, bj, wj);
      for (int i = 0; i < L->nx; i++) {
        int f = L->offset + j * L->nx + i;
        if (F->fixed[f]) continue;
        int bi[2];
        double wi[2];
        int ni = bracket(i, L->nx, bi, wi);
        int k = 0;
        double sum = 0;
        for (int a = 0; a < nj; a++) {
          int dj = coarse_to_fine(bj[a], L->ny) - j;
          for (int b = 0; b < ni; b++) {
            int di = coarse_to_fine(bi[b], L->nx) - i;
            int ok = 1;
            if (di != 0 && dj != 0) {
              int m1 = op_nbr(F, f, di, 0), m2 = op_nbr(F, f, 0, dj);
              ok = (linked(F, f, di, 0) && linked(F, m1, 0, dj)) || (linked(F, f, 0, dj) && linked(F, m2, di, 0));
            } else if (di != 0 || dj != 0) {
              ok = linked(F, f, di, dj);
            }
            P->idx[4 * f + k] = C->offset + bj[a] * C->nx + bi[b];
            P->w[4 * f + k] = ok ? wi[b] * wj[a] : 0;
            sum += P->w[4 * f + k];
            k++;
          }
        }
        for (int q = 0; q < k; q++) P->w[4 * f + q] = sum > 0 ? P->w[4 * f + q] / sum : 0;
      }
    }
  }
}

/* galerkin (newtonMultigrid.ts:345-387) */
static int galerkin(const op_t* A, const lay_t* coarse, const prol_t* P, op_t* C) {
  op_init(C, coarse, A->nl);
  for (int a = 0; a < A->N; a++) {
    if (A->fixed[a]) continue;
    for (int dj = -2; dj <= 2; dj++)
      for (int di = -2; di <= 2; di++) {
        int b = op_nbr(A, a, di, dj);
        if (b < 0 || A->fixed[b]) continue;
        const double* blk = A->A + ((size_t)a * STENCIL + offset_index(di, dj)) * 9;
        int nz = 0;
        for (int q = 0; q < 9; q++) nz |= blk[q] != 0;
        if (!nz) continue;
        for (int ka = 0; ka < 4; ka++) {
          int I = P->idx[4 * a + ka];
          if (I < 0) continue;
          double wa = P->w[4 * a + ka];
          for (int kb = 0; kb < 4; kb++) {
            int J = P->idx[4 * b + kb];
            if (J < 0) continue;
            double w = wa * P->w[4 * b + kb];
            int oi = C->ij[3 * J] - C->ij[3 * I], oj = C->ij[3 * J + 1] - C->ij[3 * I + 1];
            if (abs(oi) > 2 || abs(oj) > 2) return -1;
            add_block(C, I, offset_index(oi, oj), w, blk);
          }
        }
      }
  }
  for (int I = 0; I < C->N; I++) {
    double* base = C->A + ((size_t)I * STENCIL + CENTER) * 9;
    if (base[0] + base[4] + base[8] <= 0) {
      C->fixed[I] = 1;
      memset(C->A + (size_t)I * STENCIL * 9, 0, sizeof(double) * STENCIL * 9);
      base[0] = base[4] = base[8] = 1;
    }
  }
  return 0;
}

static void free_levels(or_solver* S, int from) {
  for (int k = from; k < S->nlev; k++) {
    op_free(&S->lv[k]);
    free(S->pr[k].idx);
    free(S->pr[k].w);
    S->pr[k].idx = 0;
    S->pr[k].w = 0;
  }
  if (from > 0 && from - 1 < S->nlev) {
    free(S->pr[from - 1].idx);
    free(S->pr[from - 1].w);
    S->pr[from - 1].idx = 0;
    S->pr[from - 1].w = 0;
  }
  if (S->nlev > from) S->nlev = from;
}

/* buildHierarchy (newtonMultigrid.ts:480-491) on the current level 0 */
/* Exact solve of the coarsest level (BD_COARSE_EXACT; an extension, the reference smooths it):
   Cholesky factor L of its DOF matrix (DOF 3 n + c, row-major nodes) within its bandwidth, FP64.
   L(i, j), i - bw <= j <= i, is stored at L[i * (bw + 1) + i - j]. Returns -1 on a non-positive
   pivot. */
static int band_bandwidth(const op_t* A) {
  int bw = 2;
  for (int l = 0; l < A->nl; l++) {
    int b = 3 * (2 * A->lay[l].nx + 2) + 2;
    if (b > bw) bw = b;
  }
  return bw;
}

static double dof_entry(const op_t* A, int i, int j) {
  int a = i / 3, r = i % 3, b = j / 3, c = j % 3;
  if (A->ij[3 * a + 2] != A->ij[3 * b + 2]) return 0;
  int di = A->ij[3 * b] - A->ij[3 * a], dj = A->ij[3 * b + 1] - A->ij[3 * a + 1];
  if (abs(di) > 2 || abs(dj) > 2) return 0;
  return A->A[((size_t)a * STENCIL + offset_index(di, dj)) * 9 + r * 3 + c];
}

static int band_cholesky(const op_t* A, int* n_out, int* bw_out, double** L_out) {
  const int n = 3 * A->N, bw = band_bandwidth(A), w = bw + 1;
  double* L = calloc((size_t)n * w, sizeof(double));
  for (int i = 0; i < n; i++) {
    int j0 = i - bw < 0 ? 0 : i - bw;
    for (int j = j0; j <= i; j++) {
      double s = dof_entry(A, i, j);
      int k0 = j0 > j - bw ? j0 : j - bw;
      for (int k = k0; k < j; k++) s -= L[(size_t)i * w + i - k] * L[(size_t)j * w + j - k];
      if (i == j) {
        if (!(s > 0)) {
          free(L);
          return -1;
        }
        L[(size_t)i * w] = sqrt(s);
      } else {
        L[(size_t)i * w + i - j] = s / L[(size_t)j * w];
      }
    }
  }
  *n_out = n;
  *bw_out = bw;
  *L_out = L;
  return 0;
}

static void band_solve(int n, int bw, const double* L, const double* b, double* x) {
  const int w = bw + 1;
  for (int i = 0; i < n; i++) {
    double s = b[i];
    int k0 = i - bw < 0 ? 0 : i - bw;
    for (int k = k0; k < i; k++) s -= L[(size_t)i * w + i - k] * x[k];
    x[i] = s / L[(size_t)i * w];
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = x[i];
    int k1 = i + bw > n - 1 ? n - 1 : i + bw;
    for (int k = i + 1; k <= k1; k++) s -= L[(size_t)k * w + k - i] * x[k];
    x[i] = s / L[(size_t)i * w];
  }
}

/* refactor: the coarsest level's exact factorization is rebuilt (always for BD_COARSE_EXACT; for
   BD_COARSE_EXACT_LAGGED only at a substep's first Newton iteration or when there is none) */
static int build_hierarchy_(or_solver* S, int refactor) {
  free_levels(S, 1);
  S->nlev = 1;
  op_t* cur = &S->lv[0];
  while (cur->N > S->P.max_coarsest_nodes && should_coarsen(cur) && S->nlev < MAXLV) {
    lay_t coarse[8];
    build_prolongation(cur, coarse, &S->pr[S->nlev - 1]);
    if (galerkin(cur, coarse, &S->pr[S->nlev - 1], &S->lv[S->nlev]) != 0) return -1;
    S->nlev++;
    cur = &S->lv[S->nlev - 1];
  }
  const int exact = S->P.coarse_solve == BD_COARSE_EXACT || S->P.coarse_solve == BD_COARSE_EXACT_LAGGED;
  if (exact && S->P.coarse_solve == BD_COARSE_EXACT_LAGGED && !refactor && S->cL) return 0;
  free(S->cL);
  S->cL = NULL;
  if (exact && band_cholesky(cur, &S->cn, &S->cbw, &S->cL) != 0) return -1;
  return 0;
}
static int build_hierarchy(or_solver* S) { return build_hierarchy_(S, 1); }

/* ---------------------------------------------------------------- line smoother */

static void inv3(const double* a, double* out) {
  double a00 = a[0], a01 = a[1], a02 = a[2], a10 = a[3], a11 = a[4], a12 = a[5], a20 = a[6], a21 = a[7], a22 = a[8];
  double c00 = a11 * a22 - a12 * a21, c01 = a12 * a20 - a10 * a22, c02 = a10 * a21 - a11 * a20;
  double det = a00 * c00 + a01 * c01 + a02 * c02;
  double id = 1 / det;
  out[0] = c00 * id;
  out[1] = (a02 * a21 - a01 * a22) * id;
  out[2] = (a01 * a12 - a02 * a11) * id;
  out[3] = c01 * id;
  out[4] = (a00 * a22 - a02 * a20) * id;
  out[5] = (a02 * a10 - a00 * a12) * id;
  out[6] = c02 * id;
  out[7] = (a01 * a20 - a00 * a21) * id;
  out[8] = (a00 * a11 - a01 * a10) * id;
}

static void pivot_inv3(const double* a, double* out, double shift) {
  double e = shift * fmax(fabs(a[0]), fmax(fabs(a[4]), fabs(a[8])));
  double s[9];
  memcpy(s, a, sizeof(s));
  s[0] += e;
  s[4] += e;
  s[8] += e;
  inv3(s, out);
}

static void mul3(const double* a, const double* b, double* out) {
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) out[r * 3 + c] = a[r * 3] * b[c] + a[r * 3 + 1] * b[3 + c] + a[r * 3 + 2] * b[6 + c];
}

/* solvePenta (newtonMultigrid.ts:435-473) */
void or_solve_penta(int n, const double* D, const double* E, const double* F, const double* r, double* x, double shift) {
  double* Dk = malloc(sizeof(double) * 9 * n);
  double* U1 = malloc(sizeof(double) * 9 * n);
  double* L1 = calloc(9 * n, sizeof(double));
  double* rk = malloc(sizeof(double) * 3 * n);
  memcpy(Dk, D, sizeof(double) * 9 * n);
  memcpy(U1, E, sizeof(double) * 9 * n);
  memcpy(rk, r, sizeof(double) * 3 * n);
  for (int k = 1; k < n; k++)
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) L1[9 * k + a * 3 + b] = E[9 * (k - 1) + b * 3 + a];
  double inv[9], Lm[9], t[9], Amk[9];
  for (int k = 0; k < n; k++) {
    pivot_inv3(Dk + 9 * k, inv, shift);
    for (int m = k + 1; m <= k + 2 && m < n; m++) {
      if (m == k + 1)
        memcpy(Amk, L1 + 9 * m, sizeof(Amk));
      else
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) Amk[a * 3 + b] = F[9 * k + b * 3 + a];
      mul3(Amk, inv, Lm);
      if (k + 1 < n) {
        mul3(Lm, U1 + 9 * k, t);
        double* dst = m == k + 1 ? Dk : L1;
        for (int q = 0; q < 9; q++) dst[9 * m + q] -= t[q];
      }
      if (k + 2 < n) {
        mul3(Lm, F + 9 * k, t);
        double* dst = m == k + 1 ? U1 : Dk;
        for (int q = 0; q < 9; q++) dst[9 * m + q] -= t[q];
      }
      for (int q = 0; q < 3; q++) rk[3 * m + q] -= Lm[q * 3] * rk[3 * k] + Lm[q * 3 + 1] * rk[3 * k + 1] + Lm[q * 3 + 2] * rk[3 * k + 2];
    }
  }
  for (int k = n - 1; k >= 0; k--) {
    double b0[3] = {rk[3 * k], rk[3 * k + 1], rk[3 * k + 2]};
    for (int q = 0; q < 3; q++) {
      if (k + 1 < n) b0[q] -= U1[9 * k + q * 3] * x[3 * (k + 1)] + U1[9 * k + q * 3 + 1] * x[3 * (k + 1) + 1] + U1[9 * k + q * 3 + 2] * x[3 * (k + 1) + 2];
      if (k + 2 < n) b0[q] -= F[9 * k + q * 3] * x[3 * (k + 2)] + F[9 * k + q * 3 + 1] * x[3 * (k + 2) + 1] + F[9 * k + q * 3 + 2] * x[3 * (k + 2) + 2];
    }
    pivot_inv3(Dk + 9 * k, inv, shift);
    for (int q = 0; q < 3; q++) x[3 * k + q] = inv[q * 3] * b0[0] + inv[q * 3 + 1] * b0[1] + inv[q * 3 + 2] * b0[2];
  }
  free(Dk);
  free(U1);
  free(L1);
  free(rk);
}

static double frob(const double* B) {
  double s = 0;
  for (int q = 0; q < 9; q++) s += B[q] * B[q];
  return sqrt(s);
}

/* smoothLines (newtonMultigrid.ts:659-733) */
static void smooth_lines(const op_t* A, double* x, const double* b, int reverse, int colors) {
  int maxLen = 0;
  for (int l = 0; l < A->nl; l++) {
    if (A->lay[l].nx > maxLen) maxLen = A->lay[l].nx;
    if (A->lay[l].ny > maxLen) maxLen = A->lay[l].ny;
  }
  double* xOld = malloc(sizeof(double) * 3 * A->N);
  int order[12][2], no = 0;
  int fams[2] = {1, 0};
  for (int fi = 0; fi < 2; fi++)
    for (int c = 0; c < colors; c++) {
      order[no][0] = fams[fi];
      order[no][1] = c;
      no++;
    }
  for (int oi = 0; oi < no; oi++) {
    int o = reverse ? no - 1 - oi : oi;
    int family = order[o][0], color = order[o][1];
    memcpy(xOld, x, sizeof(double) * 3 * A->N);
    for (int l = 0; l < A->nl; l++) {
      const lay_t* L = &A->lay[l];
      int lines = family == 0 ? L->ny : L->nx;
      int len = family == 0 ? L->nx : L->ny;
      /* the lines of one colour read the snapshot and write only their own nodes: threads change
         no result */
#pragma omp parallel
      {
        double* D = malloc(sizeof(double) * 9 * maxLen);
        double* E = malloc(sizeof(double) * 9 * maxLen);
        double* F = calloc(9 * maxLen, sizeof(double));
        double* r = malloc(sizeof(double) * 3 * maxLen);
        double* sol = calloc(3 * maxLen, sizeof(double));
#pragma omp for schedule(static)
        for (int line = color; line < lines; line += colors) {
          memset(D, 0, sizeof(double) * 9 * len);
          memset(E, 0, sizeof(double) * 9 * len);
          for (int k = 0; k < len; k++) {
            int n = family == 0 ? L->offset + line * L->nx + k : L->offset + k * L->nx + line;
            r[3 * k] = b[3 * n];
            r[3 * k + 1] = b[3 * n + 1];
            r[3 * k + 2] = b[3 * n + 2];
            if (A->fixed[n]) {
              D[9 * k] = D[9 * k + 4] = D[9 * k + 8] = 1;
              continue;
            }
            double maj = 0;
            for (int dj = -2; dj <= 2; dj++)
              for (int di = -2; di <= 2; di++) {
                int m = op_nbr(A, n, di, dj);
                if (m < 0 || (A->fixed[m] && !(di == 0 && dj == 0))) continue;
                const double* base = A->A + ((size_t)n * STENCIL + offset_index(di, dj)) * 9;
                int along = family == 0 ? di : dj;
                int across = family == 0 ? dj : di;
                if (across == 0 && along == 0) {
                  for (int q = 0; q < 9; q++) D[9 * k + q] += base[q];
                  continue;
                }
                if (across == 0 && along == 1) {
                  for (int q = 0; q < 9; q++) E[9 * k + q] = base[q];
                  continue;
                }
                if (across == 0 && along == -1) continue;
                /* colours 2 and 3 as newtonMultigrid.ts:705; 1 colour (this build's extension)
                   majorises only the along-line couplings */
                if (across == 0 || (colors > 1 && abs(across) == colors)) maj += frob(base);
                double x0 = xOld[3 * m], x1 = xOld[3 * m + 1], x2 = xOld[3 * m + 2];
                r[3 * k] -= base[0] * x0 + base[1] * x1 + base[2] * x2;
                r[3 * k + 1] -= base[3] * x0 + base[4] * x1 + base[5] * x2;
                r[3 * k + 2] -= base[6] * x0 + base[7] * x1 + base[8] * x2;
              }
            D[9 * k] += maj;
            D[9 * k + 4] += maj;
            D[9 * k + 8] += maj;
            r[3 * k] += maj * xOld[3 * n];
            r[3 * k + 1] += maj * xOld[3 * n + 1];
            r[3 * k + 2] += maj * xOld[3 * n + 2];
          }
          or_solve_penta(len, D, E, F, r, sol, PIVOT_SHIFT);
          for (int k = 0; k < len; k++) {
            int n = family == 0 ? L->offset + line * L->nx + k : L->offset + k * L->nx + line;
            if (A->fixed[n]) continue;
            x[3 * n] = sol[3 * k];
            x[3 * n + 1] = sol[3 * k + 1];
            x[3 * n + 2] = sol[3 * k + 2];
          }
        }
        free(D);
        free(E);
        free(F);
        free(r);
        free(sol);
      }
    }
  }
  free(xOld);
}

/* Threads of the OpenMP loops (op_apply, smooth_lines); returns the number in use (1 without
   OpenMP). Results do not depend on it. */
int or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* vcycle (newtonMultigrid.ts:495-528), pre = post = 1 */
/* FP16 with round to nearest even (10 stored mantissa bits; subnormals in steps of 2^-24), saturated
 * at the largest finite half, as the GPU's v_cvt_f16_f32 of the saturated value */
static double fp16_round(double a) {
  float f = (float)a;
  if (f > 65504.f) f = 65504.f;
  if (f < -65504.f) f = -65504.f;
  if (fabsf(f) < 6.103515625e-05f) return nearbyint((double)f * 16777216.0) / 16777216.0;
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0xFFFu + ((u >> 13) & 1u);
  u &= ~0x1FFFu;
  memcpy(&f, &u, 4);
  return (double)f;
}

/* BD_PRECISION_FP16 (csrc/oprd.h, k_pack_half): the V-cycle's smoother and residual read a coarse
 * level's operator rounded to FP16 at the level's scale 2^e, e such that the largest diagonal entry of a free
 * row lands in (2^13, 2^14]. Returns a rounded copy (caller frees), or NULL in FP32 mode. */
static double* half_operator(const or_solver* S, const op_t* A) {
  if (S->P.precision != BD_PRECISION_FP16 || A == &S->lv[0]) return NULL;  /* coarse levels only */
  const size_t m = (size_t)A->N * 225;
  double mx = 0;
  for (int n = 0; n < A->N; n++) {
    if (A->fixed[n]) continue;
    const double* c = A->A + (size_t)n * 225 + 12 * 9;
    mx = fmax(mx, fmax(fabs((double)(float)c[0]), fmax(fabs((double)(float)c[4]), fabs((double)(float)c[8]))));
  }
  double sc = 1.0;
  if (mx > 0) {
    int ex;
    const double fr = frexp((double)(float)mx, &ex);
    sc = ldexp(1.0, 14 - (fr == 0.5 ? ex - 1 : ex));
  }
  double* r = malloc(sizeof(double) * m);
  for (size_t q = 0; q < m; q++) r[q] = fp16_round(A->A[q] * sc) / sc;
  return r;
}

static void vcycle(or_solver* S, int k, double* x, const double* b) {
  op_t Ah = S->lv[k];
  const op_t* A = &S->lv[k];
  int colors = S->colors[k] ? S->colors[k] : 3;
  if (k == S->nlev - 1) {
    if (S->cL) {
      band_solve(S->cn, S->cbw, S->cL, b, x);
      return;
    }
    double* h = half_operator(S, A);
    if (h) {
      Ah.A = h;
      A = &Ah;
    }
    for (int s = 0; s < S->P.coarse_sweeps; s++) smooth_lines(A, x, b, s & 1, colors);
    free(h);
    return;
  }
  double* h = half_operator(S, A);
  if (h) {
    Ah.A = h;
    A = &Ah;
  }
  const prol_t* P = &S->pr[k];
  for (int sw = 0; sw < S->P.smooth_sweeps; sw++) smooth_lines(A, x, b, 0, colors);
  double* Ax = malloc(sizeof(double) * 3 * A->N);
  op_apply(A, x, Ax);
  const op_t* C = &S->lv[k + 1];
  double* rc = calloc(3 * C->N, sizeof(double));
  for (int f = 0; f < A->N; f++) {
    if (A->fixed[f]) continue;
    for (int kk = 0; kk < 4; kk++) {
      int c = P->idx[4 * f + kk];
      if (c < 0) continue;
      double w = P->w[4 * f + kk];
      for (int q = 0; q < 3; q++) rc[3 * c + q] += w * (b[3 * f + q] - Ax[3 * f + q]);
    }
  }
  for (int c = 0; c < C->N; c++)
    if (C->fixed[c]) rc[3 * c] = rc[3 * c + 1] = rc[3 * c + 2] = 0;
  double* xc = calloc(3 * C->N, sizeof(double));
  vcycle(S, k + 1, xc, rc);
  for (int f = 0; f < A->N; f++) {
    if (A->fixed[f]) continue;
    for (int kk = 0; kk < 4; kk++) {
      int c = P->idx[4 * f + kk];
      if (c < 0) continue;
      double w = P->w[4 * f + kk];
      for (int q = 0; q < 3; q++) x[3 * f + q] += w * xc[3 * c + q];
    }
  }
  for (int sw = 0; sw < S->P.smooth_sweeps; sw++) smooth_lines(A, x, b, 1, colors);
  free(Ax);
  free(rc);
  free(xc);
  free(h);
}

static double dotn(const double* a, const double* b, int n) {
  double s = 0;
  for (int q = 0; q < n; q++) s += a[q] * b[q];
  return s;
}

/* pcg (newtonMultigrid.ts:537-576); rel_tol > 0 adds the relative-residual stop (build extension) */
static void pcg(or_solver* S, const double* b, double* x, int iterations, double rel_tol, int max_iter) {
  const op_t* A = &S->lv[0];
  int n = 3 * A->N;
  double* r = malloc(sizeof(double) * n);
  double* Ap = malloc(sizeof(double) * n);
  double* z = calloc(n, sizeof(double));
  double* p = malloc(sizeof(double) * n);
  op_apply(A, x, Ap);
  for (int q = 0; q < n; q++) r[q] = b[q] - Ap[q];
  for (int m = 0; m < A->N; m++)
    if (A->fixed[m]) r[3 * m] = r[3 * m + 1] = r[3 * m + 2] = 0;
  double bn = sqrt(dotn(r, r, n));
  int limit = rel_tol > 0 ? max_iter : iterations;
  S->pcg_iters = 0;
  S->pcg_relres = bn > 0 ? 1.0 : 0.0;
  if (rel_tol > 0 && !(bn > 0)) goto done;
  vcycle(S, 0, z, r);
  memcpy(p, z, sizeof(double) * n);
  double rz = dotn(r, z, n);
  for (int it = 0; it < limit; it++) {
    op_apply(A, p, Ap);
    for (int m = 0; m < A->N; m++)
      if (A->fixed[m]) Ap[3 * m] = Ap[3 * m + 1] = Ap[3 * m + 2] = 0;
    double pAp = dotn(p, Ap, n);
    if (!(pAp > 0)) break;
    double alpha = rz / pAp;
    for (int q = 0; q < n; q++) {
      x[q] += alpha * p[q];
      r[q] -= alpha * Ap[q];
    }
    S->pcg_iters = it + 1;
    S->pcg_relres = bn > 0 ? sqrt(dotn(r, r, n)) / bn : 0;
    if (rel_tol > 0 && S->pcg_relres <= rel_tol) break;
    if (it + 1 == limit) break;
    memset(z, 0, sizeof(double) * n);
    vcycle(S, 0, z, r);
    double rz2 = dotn(r, z, n);
    double beta = rz2 / rz;
    rz = rz2;
    for (int q = 0; q < n; q++) p[q] = z[q] + beta * p[q];
  }
done:
  free(r);
  free(Ap);
  free(z);
  free(p);
}

/* chooseStep (newtonMultigrid.ts:597-603) */
double or_choose_step(const double* e, const double* m) {
  int best = 0;
  for (int k = 1; k < 4; k++)
    if (e[k] < e[best]) best = k;
  if (e[best] < -LINE_NOISE * m[best]) return LINE_ALPHAS[best];
  int all = 1;
  for (int k = 0; k < 4; k++)
    if (!(e[k] > LINE_NOISE * m[k])) all = 0;
  if (all) return 0;
  return 1;
}

/* ---------------------------------------------------------------- public solver API */

or_solver* or_create(const bd_scene* sc, const bd_params* prm) {
  or_solver* S = calloc(1, sizeof(or_solver));
  int N = sc->node_count;
  S->N = N;
  S->nl = sc->layer_count;
  for (int l = 0; l < S->nl; l++) S->lay[l] = sc->layers[l];
  S->P = *prm;
  S->C = sc->crossing_count;
#define DUP(dst, src, cnt, T) do { dst = malloc(sizeof(T) * (size_t)(cnt) + 1); memcpy(dst, src, sizeof(T) * (size_t)(cnt)); } while (0)
  DUP(S->mass, sc->mass, N, double);
  DUP(S->pinned, sc->pinned, N, uint8_t);
  DUP(S->cn0, sc->crossing_normal0, N, double);
  DUP(S->fstrength, sc->field_strength, 2 * S->C, float);
  DUP(S->farea, sc->field_area, 2 * S->C, uint8_t);
  DUP(S->damage, sc->field_damage, 2 * S->C, uint8_t);
  DUP(S->x, sc->position, 3 * N, double);
  DUP(S->rest, sc->rest_length, 2 * N, double);
  DUP(S->softening, sc->softening, 2 * N, double);
  DUP(S->intact, sc->intact, 2 * N, double);
  DUP(S->cut_mask, sc->cut_mask, 2 * N, uint32_t);
  DUP(S->fail_site, sc->fail_site, 2 * N, double);
  DUP(S->seg_flags, sc->seg_flags, 2 * N, uint16_t);
  DUP(S->tails, sc->tails, 4 * N, double);
#undef DUP
  S->v = calloc(3 * N, sizeof(double));
  S->dx = calloc(3 * N, sizeof(double));
  S->pred = calloc(3 * N, sizeof(double));
  S->strain = calloc(2 * N, sizeof(double));
  S->last_update = calloc(N, sizeof(double));
  S->nbr = malloc(sizeof(int) * 4 * N);
  S->layer_of = calloc(N, 1);
  S->grip_of = malloc(sizeof(int) * N);
  S->grip_w = calloc(N, sizeof(double));
  S->grip_off = calloc(3 * N, sizeof(double));
  for (int n = 0; n < 4 * N; n++) S->nbr[n] = -1;
  for (int n = 0; n < N; n++) S->grip_of[n] = -1;
  for (int l = 0; l < S->nl; l++) {
    const bd_layer* L = &S->lay[l];
    for (int j = 0; j < L->ny; j++)
      for (int i = 0; i < L->nx; i++) {
        int n = L->node_offset + j * L->nx + i;
        S->layer_of[n] = (uint8_t)l;
        if (i > 0) S->nbr[4 * n] = n - 1;
        if (i < L->nx - 1) S->nbr[4 * n + 1] = n + 1;
        if (j > 0) S->nbr[4 * n + 2] = n - L->nx;
        if (j < L->ny - 1) S->nbr[4 * n + 3] = n + L->nx;
      }
  }
  return S;
}

void or_destroy(or_solver* S) {
  if (!S) return;
  free_levels(S, 0);
  free(S->mass); free(S->pinned); free(S->cn0); free(S->fstrength); free(S->farea); free(S->damage);
  free(S->x); free(S->v); free(S->dx); free(S->pred); free(S->rest); free(S->softening); free(S->intact);
  free(S->cut_mask); free(S->fail_site); free(S->seg_flags); free(S->tails); free(S->strain);
  free(S->last_update); free(S->nbr); free(S->layer_of); free(S->grip_of); free(S->grip_w); free(S->grip_off);
  free(S->ext); free(S->cmeta); free(S->cL); free(S->curve);
  free(S);
}

void or_set_grips(or_solver* S, const bd_grips* G) {
  for (int n = 0; n < S->N; n++) S->grip_of[n] = -1;
  S->grip_count = 0;
  if (!G) return;
  S->grip_count = G->grip_count;
  for (int k = 0; k < G->grip_count; k++) {
    S->gk[k] = G->stiffness[k];
    for (int q = 0; q < 3; q++) S->goal[k][q] = G->goal[3 * k + q];
  }
  for (int e = 0; e < G->entry_count; e++) {
    int n = G->node[e];
    S->grip_of[n] = G->grip[e];
    S->grip_w[n] = G->weight[e];
    for (int q = 0; q < 3; q++) S->grip_off[3 * n + q] = G->offset[3 * e + q];
  }
}

/* external terms (9 per node, NULL clears), contact bounds (8 per node, NULL clears), the floor */
void or_set_ext(or_solver* S, const double* ext) {
  free(S->ext);
  S->ext = NULL;
  if (!ext) return;
  S->ext = malloc(sizeof(double) * 9 * (size_t)S->N);
  memcpy(S->ext, ext, sizeof(double) * 9 * (size_t)S->N);
}
void or_set_cmeta(or_solver* S, const double* cmeta) {
  free(S->cmeta);
  S->cmeta = NULL;
  if (!cmeta) return;
  S->cmeta = malloc(sizeof(double) * 8 * (size_t)S->N);
  memcpy(S->cmeta, cmeta, sizeof(double) * 8 * (size_t)S->N);
}
void or_set_floor(or_solver* S, int on, double z) {
  S->floor_on = on;
  S->floor_z = on ? z : 0;
}

/* projectStart (step.wgsl:85-95) */
void or_project_start(or_solver* S) {
  static const double zero[3] = {0, 0, 0};
  for (int n = 0; n < S->N; n++)
    if (!S->pinned[n]) feasible_d(S, n, zero, &S->dx[3 * n]);
}

/* predict (solver.ts:602-613) */
void or_predict(or_solver* S) {
  S->lagged = 0;
  double h = S->P.h;
  for (int n = 0; n < S->N; n++) {
    int pin = S->pinned[n];
    for (int q = 0; q < 3; q++) {
      double hv = pin ? 0 : h * S->v[3 * n + q];
      S->pred[3 * n + q] = pin ? 0 : hv + h * h * S->P.gravity[q];
      S->dx[3 * n + q] = hv;
    }
  }
}

/* fine operator at the current iterate into level 0 (keeps coarse levels when lagging) */
void or_assemble(or_solver* S) {
  op_free(&S->lv[0]);
  assemble_fine(S, &S->lv[0]);
  if (S->nlev < 1) S->nlev = 1;
}

void or_rhs(or_solver* S, double* b) {
  double g[3];
  for (int n = 0; n < S->N; n++) {
    if (S->pinned[n]) {
      b[3 * n] = b[3 * n + 1] = b[3 * n + 2] = 0;
      continue;
    }
    or_node_gradient(S, n, g);
    b[3 * n] = -g[0];
    b[3 * n + 1] = -g[1];
    b[3 * n + 2] = -g[2];
  }
}

int or_build_hierarchy(or_solver* S) { return build_hierarchy(S); }

/* replace the fine operator's blocks (e.g. by the GPU's, widened) and rebuild the coarse levels */
int or_load_fine_operator(or_solver* S, const double* A, const uint8_t* fixed) {
  if (S->nlev < 1) or_assemble(S);
  memcpy(S->lv[0].A, A, sizeof(double) * (size_t)S->N * STENCIL * 9);
  memcpy(S->lv[0].fixed, fixed, S->N);
  return build_hierarchy(S);
}

int or_level_count(or_solver* S) { return S->nlev; }
int or_level_nodes(or_solver* S, int k) { return S->lv[k].N; }
double* or_level_op(or_solver* S, int k) { return S->lv[k].A; }
uint8_t* or_level_fixed(or_solver* S, int k) { return S->lv[k].fixed; }
int* or_level_pidx(or_solver* S, int k) { return S->pr[k].idx; }
double* or_level_pw(or_solver* S, int k) { return S->pr[k].w; }

void or_apply(or_solver* S, int k, const double* x, double* y) { op_apply(&S->lv[k], x, y); }
double or_fp16_round(double a) { return fp16_round(a); }
void or_smooth(or_solver* S, int k, double* x, const double* b, int reverse) {
  op_t Ah = S->lv[k];
  double* h = half_operator(S, &S->lv[k]);
  if (h) Ah.A = h;
  smooth_lines(&Ah, x, b, reverse, S->colors[k] ? S->colors[k] : 3);
  free(h);
}
/* line colours per family of level k's smoother (1, 2 or 3) */
int or_set_line_colors(or_solver* S, int k, int colors) {
  if (k < 0 || k >= MAXLV || colors < 1 || colors > 3) return -1;
  S->colors[k] = colors;
  return 0;
}
void or_vcycle(or_solver* S, double* x, const double* b) { vcycle(S, 0, x, b); }
void or_pcg(or_solver* S, const double* b, double* x, int iterations, double rel_tol, int max_iter) { pcg(S, b, x, iterations, rel_tol, max_iter); }
int or_pcg_iterations(or_solver* S) { return S->pcg_iters; }
double or_pcg_relres(or_solver* S) { return S->pcg_relres; }

/* newtonStepPCG (newtonMultigrid.ts:610-645); returns max |component| of the applied step */
double or_newton_step(or_solver* S, int first) {
  or_assemble(S);
  int N = S->N;
  double* b = malloc(sizeof(double) * 3 * N);
  or_rhs(S, b);
  if (S->P.lag_coarse && S->lagged && !first) {
    /* keep coarse levels and the level-0 prolongation of the first iteration */
  } else {
    build_hierarchy_(S, first);
    S->lagged = 1;
  }
  double* d = calloc(3 * N, sizeof(double));
  pcg(S, b, d, S->P.pcg_iterations, S->P.pcg_rel_tol, S->P.pcg_max_iterations);
  double alpha = 1;
  if (S->P.line_search) {
    or_energy_change(S, d, S->last_e, S->last_m);
    alpha = or_choose_step(S->last_e, S->last_m);
    if (alpha == 0) S->rejections++;
  }
  S->last_alpha = alpha;
  double m = 0;
  if (S->cmeta || S->floor_on) {
    /* newtonApply (step.wgsl:97-112): the updated step made feasible for the bounds and the floor */
    for (int n = 0; n < N; n++) {
      if (S->pinned[n]) {
        d[3 * n] = d[3 * n + 1] = d[3 * n + 2] = 0;
        continue;
      }
      double old[3] = {S->dx[3 * n], S->dx[3 * n + 1], S->dx[3 * n + 2]}, nd[3];
      for (int q = 0; q < 3; q++) nd[q] = old[q] + alpha * d[3 * n + q];
      feasible_d(S, n, old, nd);
      for (int q = 0; q < 3; q++) {
        d[3 * n + q] = nd[q] - old[q];
        S->dx[3 * n + q] = nd[q];
        if (fabs(d[3 * n + q]) > m) m = fabs(d[3 * n + q]);
      }
    }
  } else {
    for (int q = 0; q < 3 * N; q++) {
      d[q] *= alpha;
      S->dx[q] += d[q];
      if (fabs(d[q]) > m) m = fabs(d[q]);
    }
  }
  for (int n = 0; n < N; n++) S->last_update[n] = norm3(&d[3 * n]);
  free(b);
  free(d);
  return m;
}

/* ---------------------------------------------------------------- yarn damage field (yarnField.ts) */

static inline int bundle_of(const bd_layer* L, int f) { return f == 1 ? L->bundle_u : L->bundle_v; }
static inline int span_len(const bd_layer* L, int f) { return f == 1 ? L->bundle_v : L->bundle_u; }
static inline double area_of(uint8_t code) { return fmax(AREA_MIN, code / 255.0); }

static int damage_code(double d) {
  if (d >= 1) return DAMAGE_CUT;
  double c = floor(d * DAMAGE_CUT);
  return c < DAMAGE_PULLED - 1 ? (int)c : DAMAGE_PULLED - 1;
}

static int span_crossing(const bd_layer* L, int i, int j, int f, int q, int p) {
  if (f == 1) return L->field_offset + (j * L->bundle_v + L->bundle_v / 2 + p) * L->field_k + i * L->bundle_u + q;
  return L->field_offset + (j * L->bundle_v + q) * L->field_k + i * L->bundle_u + L->bundle_u / 2 + p;
}

/* interpolatedStrain (yarnField.ts:216-240) */
static double interp_strain(const bd_layer* L, int f, int k, int m, const double* strain) {
  int bu = L->bundle_u, bv = L->bundle_v;
  double s = f == 1 ? (k - (bu - 1) / 2.0) / bu : (k - (bu - 0.5)) / bu;
  double t = f == 1 ? (m - (bv - 0.5)) / bv : (m - (bv - 1) / 2.0) / bv;
  double i0 = floor(s), j0 = floor(t);
  double fu = s - i0, fv = t - j0;
  int imax = f == 1 ? L->nx - 1 : L->nx - 2;
  int jmax = f == 1 ? L->ny - 2 : L->ny - 1;
  double num = 0, den = 0;
  for (int dj = 0; dj < 2; dj++)
    for (int di = 0; di < 2; di++) {
      int i = (int)i0 + di, j = (int)j0 + dj;
      if (i < 0 || j < 0 || i > imax || j > jmax) continue;
      double e = strain[2 * (L->node_offset + j * L->nx + i) + f];
      if (!isfinite(e)) continue;
      double w = (di ? fu : 1 - fu) * (dj ? fv : 1 - fv);
      num += w * e;
      den += w;
    }
  return den > 1e-6 ? num / den : NAN;
}

static int share_reach(const bd_layer* L, int f) {
  double r = SHARE_REACH / (f == 1 ? L->weave_pitch_v : L->weave_pitch_u);
  int v = (int)floor(r + 0.5);
  return v > 1 ? v : 1;
}

static double share_cap(int reach, double mu, double n0, double wrap, double tCross, double peak) {
  return ((2 * reach + 1) * mu * (n0 + wrap * tCross)) / peak;
}

/* shareFactor (yarnField.ts:303-321) */
static double share_factor(const bd_layer* L, const uint8_t* sep, int f, int k, int m, int reach, double cap) {
  int K = L->field_k, M = L->field_m;
  int along = f == 1 ? m : k, lenAlong = f == 1 ? M : K;
  int across = f == 1 ? k : m, lenAcross = f == 1 ? K : M;
  int runs = 0;
  for (int dir = -1; dir <= 1; dir += 2)
    for (int d = 1; d <= SHARE_MAXRUN; d++) {
      int y = across + dir * d;
      if (y < 0 || y >= lenAcross) break;
      int broken = 0;
      int a0 = along - reach < 0 ? 0 : along - reach;
      int a1 = along + reach > lenAlong - 1 ? lenAlong - 1 : along + reach;
      for (int a = a0; a <= a1 && !broken; a++) {
        int at = L->field_offset + (f == 1 ? a * K + y : y * K + a);
        broken = sep[2 * at + f] == 1;
      }
      if (!broken) break;
      runs++;
    }
  return 1 + fmin((SHARE_FRACTION * runs) / 2, cap);
}

/* updateField (yarnField.ts:338-410) */
static void update_field(or_solver* S) {
  int C2 = 2 * S->C;
  uint8_t* sep = malloc(C2 + 1);
  for (int q = 0; q < C2; q++) sep[q] = S->damage[q] >= DAMAGE_SEPARATED;
  for (int l = 0; l < S->nl; l++) {
    const bd_layer* L = &S->lay[l];
    const bd_material* mt = &L->material;
    int K = L->field_k;
    for (int j = 0; j < L->ny; j++)
      for (int i = 0; i < L->nx; i++)
        for (int f = 0; f < 2; f++) {
          if ((f == 0 && i >= L->nx - 1) || (f == 1 && j >= L->ny - 1)) continue;
          int s = 2 * (L->node_offset + j * L->nx + i) + f;
          if (!seg_active(S, s)) continue;
          const bd_yarn* y = &mt->yarns[f];
          int B = bundle_of(L, f), P = span_len(L, f);
          double soft = S->softening[s] / S->rest[s];
          int reach = share_reach(L, f);
          for (int q = 0; q < B; q++) {
            if ((S->cut_mask[s] >> q) & 1) continue;
            int mid = span_crossing(L, i, j, f, q, P >> 1) - L->field_offset;
            double eps = interp_strain(L, f, mid % K, mid / K, S->strain);
            if (!isfinite(eps)) continue;
            double minPeak = INFINITY;
            for (int p = 0; p < P; p++) {
              int c = span_crossing(L, i, j, f, q, p);
              double pk = or_strain_at_tension(y, ((double)S->fstrength[2 * c + f] * B) / area_of(S->farea[2 * c + f]));
              if (pk < minPeak) minPeak = pk;
            }
            if (eps * (1 + SHARE_FRACTION * SHARE_MAXRUN) <= minPeak) continue;
            int g = 1 - f;
            int kM = mid % K, mM = mid / K;
            double ec = interp_strain(L, g, kM, mM, S->strain);
            double tCross = isfinite(ec) ? (fmax(0, or_axial_tension(&mt->yarns[g], ec)) * area_of(S->farea[2 * (mid + L->field_offset) + g])) / bundle_of(L, g) : 0;
            double transfer = share_cap(reach, mt->crossing_friction, S->cn0[s >> 1] / (L->bundle_u * L->bundle_v), mt->crossing_wrap, tCross, 1);
            int best = -1, soft0 = -1, softCode = 0;
            double bestScore = -INFINITY, bestOmega = 1;
            for (int p = 0; p < P; p++) {
              int c = span_crossing(L, i, j, f, q, p);
              int local = c - L->field_offset;
              double omega = share_factor(L, sep, f, local % K, local / K, reach, transfer / S->fstrength[2 * c + f]);
              double peak = or_strain_at_tension(y, ((double)S->fstrength[2 * c + f] * B) / area_of(S->farea[2 * c + f]));
              double score = omega / peak;
              if (score > bestScore) {
                best = c;
                bestScore = score;
                bestOmega = omega;
              }
              if (S->damage[2 * c + f] > softCode) {
                soft0 = c;
                softCode = S->damage[2 * c + f];
              }
            }
            if (soft0 >= 0) {
              best = soft0;
              int local = best - L->field_offset;
              bestOmega = share_factor(L, sep, f, local % K, local / K, reach, transfer / S->fstrength[2 * best + f]);
            }
            int code = damage_code(or_yarn_damage(y, B, area_of(S->farea[2 * best + f]), S->fstrength[2 * best + f], soft, eps * bestOmega));
            if (code > S->damage[2 * best + f]) S->damage[2 * best + f] = (uint8_t)code;
          }
        }
  }
  free(sep);
}

static int break_distance(const bd_layer* L, const uint8_t* damage, int f, int k, int m) {
  int pos = f == 0 ? k : m, len = f == 0 ? L->field_k : L->field_m, stride = f == 0 ? 1 : L->field_k;
  int c = L->field_offset + m * L->field_k + k;
  int best = -1;
  for (int dir = -1; dir <= 1; dir += 2)
    for (int d = 1; d <= RECOIL_MAX; d++) {
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

double or_recoil_length(const bd_yarn* y, double B, double area, double peak, double eps, double mu, double n0, double wrap, double tCross, double pitch) {
  if (!(eps > 0)) return 0;
  double share = area / B;
  double e = fmin(eps, or_strain_at_tension(y, peak / share));
  double comp = share * (or_axial_tension(y, e) * e - or_axial_energy_density(y, e));
  return (comp * pitch) / (mu * (n0 + wrap * fmax(0, tCross)));
}

/* pullOut (yarnField.ts:125-152); returns crossings marked */
static int pull_out(or_solver* S) {
  int C2 = 2 * S->C, marked = 0;
  uint8_t* mark = calloc(C2 + 1, 1);
  for (int l = 0; l < S->nl; l++) {
    const bd_layer* L = &S->lay[l];
    const bd_material* mt = &L->material;
    for (int m = 0; m < L->field_m; m++)
      for (int k = 0; k < L->field_k; k++) {
        int c = L->field_offset + m * L->field_k + k;
        for (int f = 0; f < 2; f++) {
          if (S->damage[2 * c + f] >= DAMAGE_SEPARATED) continue;
          int d = break_distance(L, S->damage, f, k, m);
          if (d < 0) continue;
          double eps = interp_strain(L, f, k, m, S->strain);
          if (!isfinite(eps)) continue;
          int g = 1 - f;
          double ec = interp_strain(L, g, k, m, S->strain);
          double tCross = isfinite(ec) ? (fmax(0, or_axial_tension(&mt->yarns[g], ec)) * area_of(S->farea[2 * c + g])) / bundle_of(L, g) : 0;
          int i = k / L->bundle_u < L->nx - 1 ? k / L->bundle_u : L->nx - 1;
          int j = m / L->bundle_v < L->ny - 1 ? m / L->bundle_v : L->ny - 1;
          int n = L->node_offset + j * L->nx + i;
          double pitch = f == 1 ? L->weave_pitch_v : L->weave_pitch_u;
          double delta = or_recoil_length(&mt->yarns[f], bundle_of(L, f), area_of(S->farea[2 * c + f]), S->fstrength[2 * c + f], eps,
                                          mt->crossing_friction, S->cn0[n] / (L->bundle_u * L->bundle_v), mt->crossing_wrap, tCross, pitch);
          if (delta >= (d - 0.5) * pitch) {
            mark[2 * c + f] = 1;
            marked++;
          }
        }
      }
  }
  for (int q = 0; q < C2; q++)
    if (mark[q]) S->damage[q] = DAMAGE_PULLED;
  free(mark);
  return marked;
}

static void part_span_yarns(const bd_layer* L, uint8_t* damage, int i, int j, int f, int slipEnd) {
  int P = span_len(L, f);
  int half = P - P / 2;
  for (int q = 0; q < bundle_of(L, f); q++) {
    int ofA = slipEnd != 0;
    for (int p = 0; p < P; p++) {
      int e = 2 * span_crossing(L, i, j, f, q, p) + f;
      if (damage[e] >= DAMAGE_SEPARATED) {
        if (slipEnd < 0) ofA = 0;
      } else if (ofA != (p < half)) {
        damage[e] = DAMAGE_PULLED;
      }
    }
  }
}

/* coupleSpans (yarnField.ts:443-499); returns ruptures */
static int couple_spans(or_solver* S) {
  int ruptures = 0;
  for (int l = 0; l < S->nl; l++) {
    const bd_layer* L = &S->lay[l];
    for (int j = 0; j < L->ny; j++)
      for (int i = 0; i < L->nx; i++)
        for (int f = 0; f < 2; f++) {
          if ((f == 0 && i >= L->nx - 1) || (f == 1 && j >= L->ny - 1)) continue;
          int n = L->node_offset + j * L->nx + i;
          int s = 2 * n + f;
          if (S->seg_flags[s] & SEG_SLIPPED) {
            part_span_yarns(L, S->damage, i, j, f, S->fail_site[s] < 0.5 ? 0 : 1);
            S->seg_flags[s] &= (uint16_t)~SEG_SLIPPED;
            continue;
          }
          if (!seg_active(S, s)) continue;
          int B = bundle_of(L, f), P = span_len(L, f);
          double sum = 0, site = 0;
          uint32_t mask = 0;
          for (int q = 0; q < B; q++) {
            int dmax = 0, first = -1;
            double compliance = 0;
            for (int p = 0; p < P; p++) {
              int c = span_crossing(L, i, j, f, q, p);
              int d = S->damage[2 * c + f];
              if (d > dmax) dmax = d;
              if (d >= DAMAGE_SEPARATED && first < 0) first = p;
              compliance += 1 / area_of(S->farea[2 * c + f]);
            }
            sum += (dmax >= DAMAGE_SEPARATED ? 0 : 1 - (double)dmax / DAMAGE_CUT) * (P / compliance);
            if (dmax >= DAMAGE_SEPARATED) {
              mask |= 1u << q;
              site += (first + 0.5) / P;
            }
          }
          S->intact[s] = sum / B;
          S->cut_mask[s] = mask;
          uint32_t full = B == 32 ? 0xffffffffu : ((1u << B) - 1u);
          if (mask == full) {
            double phi = site / B;
            double rest = S->rest[s];
            int b = f == 0 ? n + 1 : n + L->nx;
            S->seg_flags[s] |= SEG_BROKEN;
            S->fail_site[s] = phi;
            S->tails[4 * n + 2 * f + 1] += rest * phi;
            S->tails[4 * b + 2 * f] += rest * (1 - phi);
            ruptures++;
            part_span_yarns(L, S->damage, i, j, f, -1);
          }
        }
  }
  return ruptures;
}

/* updateDamage (solver.ts:659-666) */
void or_update_damage(or_solver* S) {
  for (int q = 0; q < 2 * S->N; q++) S->strain[q] = seg_active(S, q) ? seg_length(S, q) / S->rest[q] - 1 : NAN;
  update_field(S);
  pull_out(S);
  S->broken += couple_spans(S);
}

int or_couple_spans(or_solver* S) { return couple_spans(S); }
int or_pull_out(or_solver* S) { return pull_out(S); }
void or_update_field(or_solver* S) { update_field(S); }

/* segTension (solver.ts:647-652) */
static void seg_tension(const or_solver* S, int s, double* T, double* dTdL) {
  double L = S->rest[s], l = seg_length(S, s), K;
  span_response(S, s, l / L - 1, T, &K);
  *dTdL = fmax(K, 0) * (l / (L * L));
}

/* slipNode (solver.ts:685-740) */
static void slip_node(or_solver* S, int n, int f) {
  if (S->pinned[n]) return;
  int sL = seg_of(S, n, f, -1), sR = seg_of(S, n, f, 1);
  if (sL < 0 || sR < 0) return;
  int actL = seg_active(S, sL), actR = seg_active(S, sR);
  if (!actL && !actR) return;
  int tailL = 4 * n + 2 * f, tailR = 4 * n + 2 * f + 1;
  double TL = 0, kL = 0, TR = 0, kR = 0;
  if (actL) seg_tension(S, sL, &TL, &kL);
  if (actR) seg_tension(S, sR, &TR, &kR);
  double tc = 0;
  int cnt = 0;
  for (int dir = -1; dir <= 1; dir += 2) {
    int s = seg_of(S, n, 1 - f, dir);
    if (seg_active(S, s)) {
      double T, k;
      seg_tension(S, s, &T, &k);
      tc += fmax(0, T);
      cnt++;
    }
  }
  if (cnt == 0) return;
  tc /= cnt;
  const bd_material* mat = mat_of(S, n);
  double Fc = mat->crossing_friction * (S->cn0[n] + mat->crossing_wrap * mat->yarns[f].yarns_per_span * tc);
  double dT = fmax(TR, 0) - fmax(TL, 0);
  if (fabs(dT) <= Fc) return;
  int highR = dT > 0;
  double kHigh = highR ? kR : kL;
  int lowActive = highR ? actL : actR;
  double kLow = lowActive ? (highR ? kL : kR) : 0;
  double dl = (fabs(dT) - Fc) / fmax(kHigh + kLow, 1e-9);
  double pitch = f == 0 ? mat->pitch_u : mat->pitch_v;
  int sHigh = highR ? sR : sL, sLow = highR ? sL : sR;
  if (lowActive) {
    dl = fmin(dl, fmax(0, S->rest[sLow] - 0.55 * pitch));
    S->rest[sLow] -= dl;
    S->rest[sHigh] += dl;
  } else {
    int tIdx = highR ? tailL : tailR;
    dl = fmin(dl, S->tails[tIdx]);
    S->tails[tIdx] -= dl;
    S->rest[sHigh] += dl;
    if (S->tails[tIdx] <= 1e-9) {
      S->tails[tIdx] = 0;
      int far = highR ? seg_other(S, sHigh) : sHigh >> 1;
      int farTail = highR ? 4 * far + 2 * f : 4 * far + 2 * f + 1;
      S->tails[farTail] += S->rest[sHigh];
      S->seg_flags[sHigh] |= SEG_BROKEN | SEG_SLIPPED;
      S->fail_site[sHigh] = highR ? 0 : 1;
      S->pulled++;
    }
  }
}

/* updateSlip (solver.ts:669-683) */
void or_update_slip(or_solver* S) {
  for (int f = 0; f < 2; f++)
    for (int parity = 0; parity < 2; parity++)
      for (int l = 0; l < S->nl; l++) {
        const bd_layer* L = &S->lay[l];
        for (int j = 0; j < L->ny; j++)
          for (int i = 0; i < L->nx; i++) {
            int along = f == 0 ? i : j;
            if ((along & 1) != parity) continue;
            slip_node(S, L->node_offset + j * L->nx + i, f);
          }
      }
}

/* finalize (solver.ts:622-632) */
void or_finalize(or_solver* S) {
  double h = S->P.h;
  for (int q = 0; q < 3 * S->N; q++) {
    S->v[q] = S->dx[q] / h;
    S->x[q] += S->dx[q];
    S->dx[q] = 0;
  }
  if (!S->P.history) return;
  or_update_damage(S);
  or_update_slip(S);
}

/* substep (solver.ts:615-619) with the gated optional iterations of encodeSubstep (gpu/solver.ts:737-760) */
int or_substep(or_solver* S) {
  or_predict(S);
  int it = 0;
  double moved = INFINITY;
  int nmax = S->P.newton_max > S->P.newton_iterations ? S->P.newton_max : S->P.newton_iterations;
  for (; it < nmax; it++) {
    if (it >= S->P.newton_iterations && !(moved > S->P.newton_tolerance)) break;
    or_newton_step(S, it == 0);
    moved = 0;
    for (int n = 0; n < S->N; n++)
      if (S->last_update[n] > moved) moved = S->last_update[n];
  }
  or_finalize(S);
  return it;
}

double or_yarn_material_length(or_solver* S) {
  double total = 0;
  for (int s = 0; s < 2 * S->N; s++)
    if (seg_active(S, s)) total += S->rest[s];
  for (int q = 0; q < 4 * S->N; q++) total += S->tails[q];
  return total;
}

/* raw state access for the Python wrapper */
void* or_ptr(or_solver* S, int which) {
  switch (which) {
    case 0: return S->x;
    case 1: return S->v;
    case 2: return S->dx;
    case 3: return S->pred;
    case 4: return S->rest;
    case 5: return S->intact;
    case 6: return S->cut_mask;
    case 7: return S->seg_flags;
    case 8: return S->fail_site;
    case 9: return S->tails;
    case 10: return S->damage;
    case 11: return S->last_update;
    case 12: return S->strain;
    case 13: return S->last_e;
    case 14: return S->last_m;
    case 15: return S->pinned;
    default: return 0;
  }
}

int or_counter(or_solver* S, int which) {
  switch (which) {
    case 0: return S->broken;
    case 1: return S->pulled;
    case 2: return S->rejections;
    default: return 0;
  }
}
double or_last_alpha(or_solver* S) { return S->last_alpha; }
void or_set_params(or_solver* S, const bd_params* p) { S->P = *p; }

/* direct access to yarn-field helpers for the fixture tests */
double or_interpolated_strain(or_solver* S, int l, int f, int k, int m) { return interp_strain(&S->lay[l], f, k, m, S->strain); }
double or_share_factor(or_solver* S, int l, const uint8_t* sep, int f, int k, int m, int reach, double cap) {
  return share_factor(&S->lay[l], sep, f, k, m, reach, cap);
}
int or_span_crossing(or_solver* S, int l, int i, int j, int f, int q, int p) { return span_crossing(&S->lay[l], i, j, f, q, p); }
int or_share_reach(or_solver* S, int l, int f) { return share_reach(&S->lay[l], f); }
double or_share_cap(int reach, double mu, double n0, double wrap, double tCross, double peak) { return share_cap(reach, mu, n0, wrap, tCross, peak); }

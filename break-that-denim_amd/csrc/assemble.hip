// Prediction, fine Newton-row assembly, Newton update and commit.
//
// assemble: one thread per lattice node computes its row of the clamp-projected Hessian of the
// incremental potential (inertia, grip, air drag, axial spans, fold-stiffening bending triples,
// shear diagonals) in the half-stencil layout of bd_internal.h, and the Newton rhs -grad Phi.
// It restates assembleFineOperator (src/simulation/reference/newtonMultigrid.ts:169-274) and
// nodeGradient (src/simulation/reference/solver.ts:387-422) as a per-row gather: every element
// touching the node is evaluated by the node's own thread, so there are no atomics and the row is
// written once with coalesced plane stores. Relative vectors (x_b - x_a) + (dx_b - dx_a) keep the
// FP32 differences exact for nearby nodes.
#include <stdexcept>

#include "bd_internal.h"
#include "bounds.h"

namespace bd {

namespace {

struct NodeCtx {
  const float *x, *dx, *pred;
  const float* mass;
  const double* cn0;
  const uint8_t *pinned;
  const double *rest, *intact;
  const uint32_t* cut;
  const uint16_t* flags;
  int Np;
};

__device__ __forceinline__ float3 ld3(const float* p, int Np, int n) { return make_float3(p[n], p[Np + n], p[2 * Np + n]); }
__device__ __forceinline__ void st3(float* p, int Np, int n, float a, float b, float c) {
  p[n] = a;
  p[Np + n] = b;
  p[2 * Np + n] = c;
}

__device__ __forceinline__ bool active(const NodeCtx& c, int s) { return s >= 0 && (c.flags[s] & (SEG_EXISTS | SEG_BROKEN)) == SEG_EXISTS; }

// (x_b + dx_b) - (x_a + dx_a), formed from exact coordinate differences
__device__ __forceinline__ void rel(const NodeCtx& c, int a, int b, float o[3]) {
  float3 xa = ld3(c.x, c.Np, a), xb = ld3(c.x, c.Np, b);
  float3 da = ld3(c.dx, c.Np, a), db = ld3(c.dx, c.Np, b);
  o[0] = (xb.x - xa.x) + (db.x - da.x);
  o[1] = (xb.y - xa.y) + (db.y - da.y);
  o[2] = (xb.z - xa.z) + (db.z - da.z);
}

struct LayerPos {
  int l, i, j, nx, ny, off;
};

__device__ __forceinline__ LayerPos locate(const Geo& g, int n) {
  LayerPos p;
  p.l = geo_layer(g, n);
  p.nx = g.nx[p.l];
  p.ny = g.ny[p.l];
  p.off = g.off[p.l];
  int r = n - p.off;
  p.j = r / p.nx;
  p.i = r - p.j * p.nx;
  return p;
}

// neighbour along family f (0: u / i, 1: v / j) in direction dir, or -1
__device__ __forceinline__ int nbr(const LayerPos& p, int n, int f, int dir) {
  if (f == 0) {
    int i = p.i + dir;
    return (i < 0 || i >= p.nx) ? -1 : n + dir;
  }
  int j = p.j + dir;
  return (j < 0 || j >= p.ny) ? -1 : n + dir * p.nx;
}

// axial span s from owner a to b: force fs along unit e (a -> b) and projected stiffnesses
// (solver.ts:254-277 segForce; newtonMultigrid.ts:192-212)
__device__ void span_axial(const NodeCtx& c, const DevLayer& L, float h, int s, int a, int b, float e[3], float& fs,
                           float& kpar, float& kperp) {
  float3 xa = ld3(c.x, c.Np, a), xb = ld3(c.x, c.Np, b);
  float3 da = ld3(c.dx, c.Np, a), db = ld3(c.dx, c.Np, b);
  float en0 = xb.x - xa.x, en1 = xb.y - xa.y, en2 = xb.z - xa.z;
  float e0 = en0 + (db.x - da.x), e1 = en1 + (db.y - da.y), e2 = en2 + (db.z - da.z);
  float l = sqrtf(e0 * e0 + e1 * e1 + e2 * e2);
  float ln = sqrtf(en0 * en0 + en1 * en1 + en2 * en2);
  e[0] = e0 / l;
  e[1] = e1 / l;
  e[2] = e2 / l;
  float Lr = (float)c.rest[s];
  float it = (float)c.intact[s];
  const DevYarn& y = L.y[s & 1];
  float eps = l / Lr - 1.f;
  float T = it * axial_tension<float>(y, eps);
  float K = it * axial_stiffness<float>(y, eps);
  float cv = (float)y.visc / (Lr * h);
  fs = T + cv * (l - ln);
  kpar = fmaxf(K, 0.f) / Lr + cv;
  kperp = fmaxf(fs, 0.f) / l;
}

// bending triple centred at b along family f (solver.ts:284-352, newtonMultigrid.ts:214-246):
// returns false if inactive; else r, wa, wc, Hessian scalar k and gradient coefficient kv
__device__ bool bend_triple(const NodeCtx& c, const DevLayer& L, const LayerPos& pb, float h, int b, int f, int& a, int& cc,
                            float r[3], float& wa, float& wc, float& k, float& kv) {
  a = nbr(pb, b, f, -1);
  cc = nbr(pb, b, f, 1);
  if (a < 0 || cc < 0) return false;
  int s0 = 2 * a + f, s1 = 2 * b + f;
  if (!active(c, s0) || !active(c, s1)) return false;
  const DevYarn& y = L.y[f];
  float kperb, fold;
  bend_geometry<float>((float)c.rest[s0], (float)c.rest[s1], wa, wc, kperb, fold);
  float sc = bend_scale<float>((float)c.intact[s0], (float)c.intact[s1], c.cut[s0], c.cut[s1], y.B);
  float kb = (float)y.bend_rig * sc * kperb;
  float kd = (kb * (float)y.bend_visc) / h;
  float ra[3], rc[3];
  rel(c, b, a, ra);
  rel(c, b, cc, rc);
  float3 xa = ld3(c.x, c.Np, a), xb = ld3(c.x, c.Np, b), xc = ld3(c.x, c.Np, cc);
  float rn0 = wa * (xa.x - xb.x) + wc * (xc.x - xb.x);
  float rn1 = wa * (xa.y - xb.y) + wc * (xc.y - xb.y);
  float rn2 = wa * (xa.z - xb.z) + wc * (xc.z - xb.z);
  for (int q = 0; q < 3; q++) r[q] = wa * ra[q] + wc * rc[q];
  float q2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  float p0, p1, p2;
  bend_psi<float>(q2, fold, p0, p1, p2);
  k = kb * (p1 + 2.f * p2 * q2) + kd;
  float lr = sqrtf(q2);
  float lrn = sqrtf(rn0 * rn0 + rn1 * rn1 + rn2 * rn2);
  kv = kb * p1 + (kd * (lr - lrn)) / fmaxf(lr, 1e-12f);
  return true;
}

// shear diagonal d of the cell with lower-left node cell (solver.ts:232-241, 374-380)
__device__ bool diag_active(const NodeCtx& c, int cell, int nx, int d) {
  int r = cell + 1, t = cell + nx;
  bool cu = active(c, 2 * cell), cv = active(c, 2 * cell + 1), rv = active(c, 2 * r + 1), tu = active(c, 2 * t);
  return d == 0 ? ((cu && rv) || (cv && tu)) : ((cu && cv) || (rv && tu));
}

__device__ void cell_shear(const NodeCtx& c, const DevLayer& L, int cell, int nx, float& ks, float& jam, float& Ld) {
  int r = cell + 1, t = cell + nx;
  auto io = [&](int s) { return active(c, s) ? (float)c.intact[s] : 0.f; };
  ks = ((float)L.shear_k * c.cn0[cell]) / (float)L.cn0;
  jam = shear_jam<float>(cell_area<float>(io(2 * cell), io(2 * t), io(2 * cell + 1), io(2 * r + 1)));
  float a = (float)c.rest[2 * cell], b = (float)c.rest[2 * cell + 1];
  Ld = sqrtf(a * a + b * b);
}

// shear diagonal p -> q: force f along unit u and its projected block (axial form)
__device__ void shear_diag(const NodeCtx& c, float ks, float jam, float Ld, int p, int q, float u[3], float& f, float& kpar,
                           float& kperp) {
  rel(c, p, q, u);
  float l = sqrtf(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  u[0] /= l;
  u[1] /= l;
  u[2] /= l;
  float k, e;
  shear_law<float>(ks, Ld, l, jam, f, k, e);
  kpar = k;
  kperp = fmaxf(f, 0.f) / l;
}

struct AsmArgs {
  NodeCtx c;
  const DevLayer* layers;
  Geo geo;
  int N;
  float h;
  float* op;
  float* rhs;
  const int* grip_of;
  const float* grip_w;
  const double* grip_off;
  const double* grip_goal_k;
  const float* ext;      // external terms (self-contact) or null
  const float4* cmeta;   // contact bounds or null
  int floor_on;
  float floor_z;
  // detached pieces: node state and the per-node blocks of the rigid-mode correction (or null)
  const uint8_t* state;
  float* rigidk;
};

// rotationOfSpring (assemble.wgsl:20-27): a rigid rotation w moves the ends of a term's vector d by
// w x d; a spring block kPar e e^T + kPerp (I - e e^T) along d resists it with kPerp (|d|^2 I - d d^T),
// a scalar stencil k on a vector r with k (|r|^2 I - r r^T)
__device__ __forceinline__ void add_rotation(Sym& S, float kperp, const float d[3]) {
  float l2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  S.a[0] += kperp * (l2 - d[0] * d[0]);
  S.a[1] -= kperp * (d[0] * d[1]);
  S.a[2] -= kperp * (d[0] * d[2]);
  S.a[3] += kperp * (l2 - d[1] * d[1]);
  S.a[4] -= kperp * (d[1] * d[2]);
  S.a[5] += kperp * (l2 - d[2] * d[2]);
}

__device__ __forceinline__ void add_neg_axial(float* blk, float kpar, float kperp, const float e[3]) {
  float d = kpar - kperp;
  blk[0] -= d * e[0] * e[0] + kperp;
  blk[1] -= d * e[0] * e[1];
  blk[2] -= d * e[0] * e[2];
  blk[3] -= d * e[1] * e[1] + kperp;
  blk[4] -= d * e[1] * e[2];
  blk[5] -= d * e[2] * e[2] + kperp;
}

__global__ __launch_bounds__(256) void k_assemble(AsmArgs A) {
  int n = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= A.N) return;
  const NodeCtx& c = A.c;
  const int Np = c.Np;
  float* op = A.op;
  if (c.pinned[n]) {
    for (int q = 0; q < kFineComps; q++) op[q * Np + n] = 0.f;
    op[0 * Np + n] = 1.f;
    op[3 * Np + n] = 1.f;
    op[5 * Np + n] = 1.f;
    st3(A.rhs, Np, n, 0.f, 0.f, 0.f);
    return;
  }
  LayerPos P = locate(A.geo, n);
  const DevLayer& L = A.layers[P.l];
  const float h = A.h;
  Sym C;
  sym_zero(C);
  float g[3] = {0.f, 0.f, 0.f};
  float blk[2][6] = {{0}};  // E (f = 0), N (f = 1)
  float ne[6] = {0}, nw[6] = {0};
  float scal2[2] = {0.f, 0.f};  // EE, NN

  // inertia (newtonMultigrid.ts:184-188, solver.ts:391-392)
  float w = c.mass[n] / (h * h);
  sym_add_diag(C, w);
  // a detached node's external block and the rotation stiffness of the terms it owns
  // (assemble.wgsl:174-176, 323-327): the rigid-mode correction's E is summed from them
  const bool detached = A.state && A.state[n] == kStateDetached;
  Sym Kx, Kr;
  sym_zero(Kx);
  sym_zero(Kr);
  sym_add_diag(Kx, w);
  float3 dn = ld3(c.dx, Np, n);
  float3 pn = ld3(c.pred, Np, n);
  g[0] += w * (dn.x - pn.x);
  g[1] += w * (dn.y - pn.y);
  g[2] += w * (dn.z - pn.z);

  // grip spring (solver.ts:173-180)
  int gi = A.grip_of[n];
  if (gi >= 0) {
    const double* G = A.grip_goal_k + 4 * gi;
    double kg = G[3] * (double)A.grip_w[n];
    float3 xn = ld3(c.x, Np, n);
    double xs[3] = {xn.x, xn.y, xn.z}, ds[3] = {dn.x, dn.y, dn.z};
    for (int q = 0; q < 3; q++) g[q] += (float)(kg * ((xs[q] - G[q]) + ds[q] - A.grip_off[q * Np + n]));
    sym_add_diag(C, (float)kg);
    sym_add_diag(Kx, (float)kg);
  }

  // implicit air drag (solver.ts:132-165), from the accepted positions
  if (L.air_drag > 0) {
    float k = (float)(L.air_drag * L.pu * L.pv) / h;
    float t[2][3], ll[2];
    float3 xn = ld3(c.x, Np, n);
    for (int f = 0; f < 2; f++) {
      int pv = nbr(P, n, f, -1), nx_ = nbr(P, n, f, 1);
      int a = (pv >= 0 && active(c, 2 * pv + f)) ? pv : n;
      int b = (nx_ >= 0 && active(c, 2 * n + f)) ? nx_ : n;
      float3 xa = a == n ? xn : ld3(c.x, Np, a);
      float3 xb = b == n ? xn : ld3(c.x, Np, b);
      t[f][0] = xb.x - xa.x;
      t[f][1] = xb.y - xa.y;
      t[f][2] = xb.z - xa.z;
      ll[f] = sqrtf(t[f][0] * t[f][0] + t[f][1] * t[f][1] + t[f][2] * t[f][2]);
    }
    float e[3], kpar, kperp;
    bool done = false;
    if (ll[0] > 0.f && ll[1] > 0.f) {
      float cr[3] = {t[0][1] * t[1][2] - t[0][2] * t[1][1], t[0][2] * t[1][0] - t[0][0] * t[1][2], t[0][0] * t[1][1] - t[0][1] * t[1][0]};
      float lc = sqrtf(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
      if (lc > 1e-3f * ll[0] * ll[1]) {
        e[0] = cr[0] / lc;
        e[1] = cr[1] / lc;
        e[2] = cr[2] / lc;
        kpar = k;
        kperp = 0.f;
        done = true;
      }
    }
    if (!done) {
      if (ll[0] > 0.f || ll[1] > 0.f) {
        int f = ll[0] >= ll[1] ? 0 : 1;
        e[0] = t[f][0] / ll[f];
        e[1] = t[f][1] / ll[f];
        e[2] = t[f][2] / ll[f];
        kpar = 0.f;
        kperp = k;
      } else {
        e[0] = 1.f;
        e[1] = 0.f;
        e[2] = 0.f;
        kpar = k;
        kperp = k;
      }
    }
    Sym D;
    sym_zero(D);
    sym_add_axial(D, 1.f, kpar, kperp, e);
    float dv[3] = {dn.x, dn.y, dn.z}, Dd[3];
    sym_mul(D.a, dv, Dd);
    for (int q = 0; q < 6; q++) {
      C.a[q] += D.a[q];
      Kx.a[q] += D.a[q];
    }
    g[0] += Dd[0];
    g[1] += Dd[1];
    g[2] += Dd[2];
  }

  // axial spans (newtonMultigrid.ts:191-212; solver.ts:396-406)
  for (int f = 0; f < 2; f++) {
    int nn = nbr(P, n, f, 1);
    if (nn >= 0 && active(c, 2 * n + f)) {
      float e[3], fs, kpar, kperp;
      span_axial(c, L, h, 2 * n + f, n, nn, e, fs, kpar, kperp);
      sym_add_axial(C, 1.f, kpar, kperp, e);
      g[0] -= fs * e[0];
      g[1] -= fs * e[1];
      g[2] -= fs * e[2];
      if (!c.pinned[nn]) add_neg_axial(blk[f], kpar, kperp, e);
      if (detached) {
        float d[3];
        rel(c, n, nn, d);
        add_rotation(Kr, kperp, d);
      }
    }
    int pp = nbr(P, n, f, -1);
    if (pp >= 0 && active(c, 2 * pp + f)) {
      float e[3], fs, kpar, kperp;
      span_axial(c, L, h, 2 * pp + f, pp, n, e, fs, kpar, kperp);
      sym_add_axial(C, 1.f, kpar, kperp, e);
      g[0] += fs * e[0];
      g[1] += fs * e[1];
      g[2] += fs * e[2];
    }
  }

  // bending triples (newtonMultigrid.ts:214-246; solver.ts:319-352)
  for (int f = 0; f < 2; f++) {
    int a, cc;
    float r[3], wa, wc, k, kv;
    if (bend_triple(c, L, P, h, n, f, a, cc, r, wa, wc, k, kv)) {  // n is the middle node
      sym_add_diag(C, 4.f * k);
      if (detached) add_rotation(Kr, k, r);
      for (int q = 0; q < 3; q++) g[q] += -2.f * kv * r[q];
      if (!c.pinned[cc]) {
        float s = -2.f * wc * k;
        blk[f][0] += s;
        blk[f][3] += s;
        blk[f][5] += s;
      }
    }
    int nn = nbr(P, n, f, 1);
    if (nn >= 0) {
      LayerPos Pn = P;
      if (f == 0) Pn.i += 1; else Pn.j += 1;
      if (bend_triple(c, L, Pn, h, nn, f, a, cc, r, wa, wc, k, kv)) {  // n is the first node
        sym_add_diag(C, wa * wa * k);
        for (int q = 0; q < 3; q++) g[q] += wa * kv * r[q];
        if (!c.pinned[nn]) {
          float s = -2.f * wa * k;
          blk[f][0] += s;
          blk[f][3] += s;
          blk[f][5] += s;
        }
        if (!c.pinned[cc]) scal2[f] += wa * wc * k;
      }
    }
    int pp = nbr(P, n, f, -1);
    if (pp >= 0) {
      LayerPos Pp = P;
      if (f == 0) Pp.i -= 1; else Pp.j -= 1;
      if (bend_triple(c, L, Pp, h, pp, f, a, cc, r, wa, wc, k, kv)) {  // n is the last node
        sym_add_diag(C, wc * wc * k);
        for (int q = 0; q < 3; q++) g[q] += wc * kv * r[q];
      }
    }
  }

  // shear diagonals of the four cells around n (newtonMultigrid.ts:247-272; solver.ts:355-421)
  {
    const int nx = P.nx;
    // cell (i, j): diagonal 0 from n to (i+1, j+1)
    if (P.i < nx - 1 && P.j < P.ny - 1) {
      int cell = n;
      if (diag_active(c, cell, nx, 0)) {
        float ks, jam, Ld, u[3], fsh, kpar, kperp;
        cell_shear(c, L, cell, nx, ks, jam, Ld);
        int q = n + nx + 1;
        shear_diag(c, ks, jam, Ld, n, q, u, fsh, kpar, kperp);
        sym_add_axial(C, 1.f, kpar, kperp, u);
        for (int z = 0; z < 3; z++) g[z] -= fsh * u[z];
        if (!c.pinned[q]) add_neg_axial(ne, kpar, kperp, u);
        if (detached) {
          float d[3];
          rel(c, n, q, d);
          add_rotation(Kr, kperp, d);
        }
      }
    }
    // cell (i-1, j): diagonal 1 from r = n to t = (i-1, j+1)
    if (P.i > 0 && P.j < P.ny - 1) {
      int cell = n - 1;
      if (diag_active(c, cell, nx, 1)) {
        float ks, jam, Ld, u[3], fsh, kpar, kperp;
        cell_shear(c, L, cell, nx, ks, jam, Ld);
        int q = n + nx - 1;
        shear_diag(c, ks, jam, Ld, n, q, u, fsh, kpar, kperp);
        sym_add_axial(C, 1.f, kpar, kperp, u);
        for (int z = 0; z < 3; z++) g[z] -= fsh * u[z];
        if (!c.pinned[q]) add_neg_axial(nw, kpar, kperp, u);
      }
    }
    // cell (i, j-1): diagonal 1 from r = (i+1, j-1) to t = n
    if (P.i < nx - 1 && P.j > 0) {
      int cell = n - nx;
      if (diag_active(c, cell, nx, 1)) {
        float ks, jam, Ld, u[3], fsh, kpar, kperp;
        cell_shear(c, L, cell, nx, ks, jam, Ld);
        int p = n - nx + 1;
        shear_diag(c, ks, jam, Ld, p, n, u, fsh, kpar, kperp);
        sym_add_axial(C, 1.f, kpar, kperp, u);
        for (int z = 0; z < 3; z++) g[z] += fsh * u[z];
        // the diagonals of cell (i, j) and, across it, of cell (i, j - 1) each count at one end
        if (detached) {
          float d[3];
          rel(c, n, p, d);
          add_rotation(Kr, kperp, d);
        }
      }
    }
    // cell (i-1, j-1): diagonal 0 from (i-1, j-1) to n
    if (P.i > 0 && P.j > 0) {
      int cell = n - nx - 1;
      if (diag_active(c, cell, nx, 0)) {
        float ks, jam, Ld, u[3], fsh, kpar, kperp;
        cell_shear(c, L, cell, nx, ks, jam, Ld);
        shear_diag(c, ks, jam, Ld, cell, n, u, fsh, kpar, kperp);
        sym_add_axial(C, 1.f, kpar, kperp, u);
        for (int z = 0; z < 3; z++) g[z] += fsh * u[z];
      }
    }
  }

  // external node terms (self-contact and friction, gathered per node): gradient and symmetric
  // Hessian (assemble.wgsl:183-190)
  if (A.ext) {
    for (int q = 0; q < 3; q++) g[q] += A.ext[q * Np + n];
    for (int q = 0; q < 6; q++) {
      C.a[q] += A.ext[(3 + q) * Np + n];
      Kx.a[q] += A.ext[(3 + q) * Np + n];
    }
  }
  // floor contact: cubic penalty along +z and friction on the step (assemble.wgsl:191-209)
  bool onFloorCore = false;
  if (A.floor_on) {
    float half = 0.5f * (float)L.thickness;
    float delta = half + kFloorBulge;
    float zc = c.x[2 * Np + n] + dn.z;
    float s = delta - (zc - A.floor_z);
    if (s > 0.f) {
      float lam = kFloorK * s * s / delta;
      float ut = sqrtf(dn.x * dn.x + dn.y * dn.y);
      float kf = kFloorMu * lam / fmaxf(ut, kFloorSpeed * h);
      g[0] += kf * dn.x;
      g[1] += kf * dn.y;
      g[2] -= lam;
      C.a[0] += kf;
      C.a[3] += kf;
      C.a[5] += 2.f * kFloorK * s / delta;
      Kx.a[0] += kf;
      Kx.a[3] += kf;
      Kx.a[5] += 2.f * kFloorK * s / delta;
      onFloorCore = zc <= A.floor_z + half * (1.f + 1e-3f);
    }
  }
  // active contact bounds: a node at its bound whose net force still drives it on is held along the
  // bound normal in the Newton direction (assemble.wgsl:305-322)
  if (A.cmeta || onFloorCore) {
    float kAct = kActiveBound * (C.a[0] + C.a[3] + C.a[5]);
    if (A.cmeta) {
      for (int b = 0; b < 2; b++) {
        float4 m = A.cmeta[2 * n + b];
        if (m.w >= 1e8f || (m.x == 0.f && m.y == 0.f && m.z == 0.f)) continue;
        float approach = -(dn.x * m.x + dn.y * m.y + dn.z * m.z);
        if (approach >= m.w - 1e-3f * fabsf(m.w) && g[0] * m.x + g[1] * m.y + g[2] * m.z > 0.f) {
          float e[3] = {m.x, m.y, m.z};
          sym_add_axial(C, 1.f, kAct, 0.f, e);
          sym_add_axial(Kx, 1.f, kAct, 0.f, e);
        }
      }
    }
    if (onFloorCore && g[2] > 0.f) {
      C.a[5] += kAct;
      Kx.a[5] += kAct;
    }
  }
  if (detached) {
    for (int q = 0; q < 6; q++) {
      A.rigidk[q * Np + n] = Kx.a[q];
      A.rigidk[(6 + q) * Np + n] = Kr.a[q];
    }
  }

  for (int q = 0; q < 6; q++) {
    op[(F_C + q) * Np + n] = C.a[q];
    op[(F_E + q) * Np + n] = blk[0][q];
    op[(F_N + q) * Np + n] = blk[1][q];
    op[(F_NE + q) * Np + n] = ne[q];
    op[(F_NW + q) * Np + n] = nw[q];
  }
  op[F_EE * Np + n] = scal2[0];
  op[F_NN * Np + n] = scal2[1];
  st3(A.rhs, Np, n, -g[0], -g[1], -g[2]);
}

// predict (solver.ts:602-613)
__global__ __launch_bounds__(256) void k_predict(int N, int Np, float h, float gx, float gy, float gz, const uint8_t* pinned,
                                                  const float* v, float* pred, float* dx) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  bool pin = pinned[n];
  float g[3] = {gx, gy, gz};
  for (int q = 0; q < 3; q++) {
    float hv = pin ? 0.f : h * v[q * Np + n];
    pred[q * Np + n] = pin ? 0.f : hv + h * h * g[q];
    dx[q * Np + n] = hv;
  }
}

// newtonApply (step.wgsl:97-112 / newtonMultigrid.ts:637-644): dx += alpha * delta with a
// non-finite firewall, per-node update norm and the iteration's largest update
struct ApplyArgs {
  int N, Np;
  const uint8_t* pinned;
  const double* delta;
  float* dx;
  float* last_update;
  const Scal* scal;
  unsigned int* max_update;
  // bounds: contact (cmeta, or null) and the floor
  const float4* cmeta;
  int floor_on;
  float floor_z, gnorm;
  const float* x;
  const float* mass;
  const uint32_t* ijl;
  const DevLayer* layers;
};

__global__ __launch_bounds__(256) void k_newton_apply(ApplyArgs A) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= A.N) return;
  const int Np = A.Np;
  if (A.pinned[n]) {
    A.last_update[n] = 0.f;
    return;
  }
  double a = A.scal->line_alpha;
  double d[3] = {A.delta[n], A.delta[Np + n], A.delta[2 * Np + n]};
  bool ok = isfinite(d[0]) && isfinite(d[1]) && isfinite(d[2]) && isfinite(a);
  float lu;
  if (!A.cmeta && !A.floor_on) {
    double s2 = 0.0;
    for (int q = 0; q < 3; q++) {
      double st = ok ? a * d[q] : 0.0;
      A.dx[q * Np + n] = (float)((double)A.dx[q * Np + n] + st);
      s2 += st * st;
    }
    lu = (float)sqrt(s2);
  } else {
    // the updated step made feasible for the contact bounds and the floor (step.wgsl:97-112)
    float old[3], nd[3];
    for (int q = 0; q < 3; q++) {
      old[q] = A.dx[q * Np + n];
      nd[q] = (float)((double)old[q] + (ok ? a * d[q] : 0.0));
    }
    if (A.cmeta) feasible_step(nd, A.cmeta[2 * n], A.cmeta[2 * n + 1]);
    if (A.floor_on) {
      float half = 0.5f * (float)A.layers[A.ijl[n] >> 24].thickness;
      nd[2] = floor_bound_z(A.x[2 * Np + n], old[2], nd[2], half, A.mass[n], A.gnorm, A.floor_z);
    }
    float e0 = nd[0] - old[0], e1 = nd[1] - old[1], e2 = nd[2] - old[2];
    lu = sqrtf(e0 * e0 + e1 * e1 + e2 * e2);
    for (int q = 0; q < 3; q++) A.dx[q * Np + n] = nd[q];
  }
  A.last_update[n] = lu;
  atomicMax(A.max_update, __float_as_uint(lu));
}

// commit (solver.ts:622-628; step.wgsl:128-147). A detached node that left the scene (below
// kRetireZ or beyond kRetireR from the axis) is retired (step.wgsl:113-126): frozen where it is, its
// velocity zeroed and its spans removed (no longer existing, not broken), so its operator row becomes
// the identity with nothing coupled to it. Two nodes retiring together clear the same flag bit.
constexpr float kRetireZ = -0.25f;
constexpr float kRetireR = 0.6f;

struct CommitArgs {
  int N, Np;
  float inv_h;
  float *x, *v, *dx;
  uint8_t* state;  // or null (no detached node)
  uint8_t* pinned;
  uint16_t* flags;
  Geo geo;
  unsigned int* max_speed;  // float bits: the largest committed speed since the host last read it
};

// commit (step.wgsl:128-147): v = dx / h, x += dx, detached debris below kRetireZ or beyond kRetireR
// retired; returns the committed speed (0 for retired nodes)
__device__ float commit_node(const CommitArgs& A, int n) {
  const int Np = A.Np;
  float d[3];
  for (int q = 0; q < 3; q++) {
    d[q] = A.dx[q * Np + n];
    A.dx[q * Np + n] = 0.f;
  }
  if (A.state) {
    const uint8_t st = A.state[n];
    if (st == kStateRetired) return 0.f;
    if (st == kStateDetached) {
      float p[3] = {A.x[n] + d[0], A.x[Np + n] + d[1], A.x[2 * Np + n] + d[2]};
      if (p[2] < kRetireZ || sqrtf(p[0] * p[0] + p[1] * p[1]) > kRetireR) {
        for (int q = 0; q < 3; q++) {
          A.x[q * Np + n] = p[q];
          A.v[q * Np + n] = 0.f;
        }
        A.state[n] = kStateRetired;
        A.pinned[n] |= kPinRetired;
        LayerPos P = locate(A.geo, n);
        for (int f = 0; f < 2; f++) {
          A.flags[2 * n + f] &= (uint16_t)~SEG_EXISTS;
          int prev = nbr(P, n, f, -1);
          if (prev >= 0) A.flags[2 * prev + f] &= (uint16_t)~SEG_EXISTS;
        }
        return 0.f;
      }
    }
  }
  for (int q = 0; q < 3; q++) {
    A.v[q * Np + n] = d[q] * A.inv_h;
    A.x[q * Np + n] += d[q];
  }
  return sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * A.inv_h;
}

// the largest speed of a wave's committed nodes into *max_speed (one atomic per wave)
__device__ __forceinline__ void wave_max_speed(unsigned int* out, float v) {
  for (int m = 32; m > 0; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, 64));
  if ((threadIdx.x & 63) == 0 && v > 0.f) atomicMax(out, __float_as_uint(v));
}

__global__ __launch_bounds__(256) void k_commit(CommitArgs A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  float speed = 0.f;
  if (n < A.N) speed = commit_node(A, n);
  wave_max_speed(A.max_speed, speed);
}

}  // namespace

void launch_predict(Solver& s) {
  ProfScope ps(s, PC_PREDICT_COMMIT, s.N * 37.0);
  k_predict<<<grid_for(s.N, 256), 256, 0, s.stream>>>(s.N, s.Np, (float)s.P.h, (float)s.P.gravity[0], (float)s.P.gravity[1],
                                                      (float)s.P.gravity[2], s.pinned, s.v, s.pred, s.dx);
}

void launch_assemble(Solver& s) {
  // reads x, dx, pred, mass, cn0, pinned, spans (rest, intact, cut, flags), grip id: 97 B;
  // writes the 32-plane row and the rhs: 140 B
  AsmArgs A;
  A.c = NodeCtx{s.x, s.dx, s.pred, s.mass, s.cn0, s.pinned, s.rest, s.intact, s.cut_mask, s.flags, s.Np};
  A.layers = s.layers;
  A.geo = s.geo;
  A.N = s.N;
  A.h = (float)s.P.h;
  A.op = s.lv[0].op;
  A.rhs = s.rhs;
  A.grip_of = s.grip_of;
  A.grip_w = s.grip_w;
  A.grip_off = s.grip_off;
  A.grip_goal_k = s.grip_goal_k;
  A.ext = (s.contact_mode != kContactNone || s.grips_active) ? s.ext : nullptr;
  A.cmeta = s.contact_mode != kContactNone ? s.cmeta : nullptr;
  A.floor_on = s.floor_on ? 1 : 0;
  A.floor_z = s.floor_z;
  A.state = s.npieces > 0 ? s.nstate : nullptr;
  A.rigidk = s.rigidk;
  {
    ProfScope ps(s, PC_ASSEMBLE, s.N * 237.0);
    k_assemble<<<grid_for(s.N, 256), 256, 0, s.stream>>>(A);
  }
  launch_colmajor_op(s, 0);
}

void launch_newton_apply(Solver& s, int it) {
  ProfScope ps(s, PC_PREDICT_COMMIT, s.N * 52.0);
  ApplyArgs A;
  A.N = s.N;
  A.Np = s.Np;
  A.pinned = s.pinned;
  A.delta = s.delta;
  A.dx = s.dx;
  A.last_update = s.last_update;
  A.scal = s.scal;
  A.max_update = &s.scal->max_update[it];
  A.cmeta = s.contact_mode != kContactNone ? s.cmeta : nullptr;
  A.floor_on = s.floor_on ? 1 : 0;
  A.floor_z = s.floor_z;
  float gx = (float)s.P.gravity[0], gy = (float)s.P.gravity[1], gz = (float)s.P.gravity[2];
  A.gnorm = sqrtf((gx * gx + gy * gy) + gz * gz);
  A.x = s.x;
  A.mass = s.mass;
  A.ijl = s.ijl;
  A.layers = s.layers;
  k_newton_apply<<<grid_for(s.N, 256), 256, 0, s.stream>>>(A);
}

void launch_commit(Solver& s) {
  ProfScope ps(s, PC_PREDICT_COMMIT, s.N * 60.0);
  CommitArgs A{s.N, s.Np, (float)(1.0 / s.P.h), s.x, s.v, s.dx, s.any_detached ? s.nstate : nullptr, s.pinned, s.flags, s.geo,
               reinterpret_cast<unsigned int*>(s.counters + 2)};
  k_commit<<<grid_for(s.N, 256), 256, 0, s.stream>>>(A);
}

}  // namespace bd

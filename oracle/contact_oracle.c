/*
 * contact_oracle.c -- scalar FP32 restatement of the reference's self-contact.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/ load this (as the checker); the product never links it. The reference's self-contact
 * exists only as WGSL (no CPU counterpart): this file restates, function by function,
 *   face hash ............ src/simulation/gpu/wgsl/selfcontact.wgsl:57-87,159-219, scan.wgsl:1-61
 *   geometry ............. selfcontact.wgsl:89-158 (closestTri, faceHitP), 221-252 (segDist, segTri, pathCrossing)
 *   exclusion window ..... selfcontact.wgsl:254-313 (rowHolds, columnHolds, latticeLinked, damageRow)
 *   detection ............ selfcontact.wgsl:315-467 (detect), 469-484 (resolveBounds)
 *   terms ................ selfcontact.wgsl:486-564 (addTerm, terms, gather)
 *   bounds ............... step.wgsl:61-95 (bound, feasible, floorBound, projectStart), common.wgsl:247-265
 * with the constants of gpu/selfcontact.ts:7-33 (SELF_CONTACT) and floor.ts (FLOOR_CONTACT).
 *
 * Scalar FP32 throughout, each sum and product rounded in the order written (compile with
 * -ffp-contract=off): sums of three terms are (x + y) + z, four-term sums left to right. The hash
 * is filled in face order; a result that depends on fill order would be a defect of the method (the
 * reference breaks every tie by face index), so comparisons sort each cell's records.
 *
 * Arrays are node-major: x, dx 3 floats per node; flags 2 per node; pairs K + 2 per node (K
 * candidates, reciprocal key, row mark); cmeta 8 per node (own bound, reciprocal bound); acc and ext
 * 9 per node; cells 2 x 2^17; items 8 floats per record (centroid now, radius, centroid at start,
 * face id as uint32 bits).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OC_K 4
#define OC_NONE 0xFFFFFFFFu
#define OC_TABLE (1u << 17)
#define OC_MASK (OC_TABLE - 1u)
#define OC_CELL 12e-3f
#define OC_REACH 3e-3f
#define OC_DELTA 1.0e-3f
#define OC_STIFF 500.f
#define OC_MU 0.45f
#define OC_MU_SPEED 0.002f
#define OC_FRACTION 0.45f
#define OC_HARD 0.3e-3f
#define OC_GS 1e8f
#define OC_HS 1e3f
#define OC_RECIP 1e-5f
#define OC_WINDOW 3
#define OC_FLOOR_BULGE 0.3e-3f
#define OC_FLOOR_K 1300.f
#define OC_UNB 1e9f

typedef struct {
  int N;
  int nl;
  int nx[8], ny[8], off[8];
  const float* x;
  const float* dx;
  const uint16_t* flags;
  const uint8_t* pinned;
  const float* mass;
} oc_scene;

static float dot3f(const float* a, const float* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static float len3f(const float* a) { return sqrtf(dot3f(a, a)); }
static void sub3f(const float* a, const float* b, float* o) {
  o[0] = a[0] - b[0];
  o[1] = a[1] - b[1];
  o[2] = a[2] - b[2];
}
static void cross3f(const float* a, const float* b, float* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

static int layer_of(const oc_scene* S, uint32_t n) {
  int l = 0;
  for (int q = 1; q < S->nl; q++)
    if ((int)n >= S->off[q]) l = q;
  return l;
}
static void node_ij(const oc_scene* S, uint32_t n, int* i, int* j, int* l) {
  *l = layer_of(S, n);
  int r = (int)n - S->off[*l];
  *j = r / S->nx[*l];
  *i = r - *j * S->nx[*l];
}
static int active(const oc_scene* S, uint32_t s) { return (S->flags[s] & 3) == 1; }
static void at_start(const oc_scene* S, uint32_t n, float* o) { memcpy(o, &S->x[3 * n], 3 * sizeof(float)); }
static void at_now(const oc_scene* S, uint32_t n, float* o) {
  for (int q = 0; q < 3; q++) o[q] = S->x[3 * n + q] + S->dx[3 * n + q];
}
static void face_corners(const oc_scene* S, uint32_t f, uint32_t c[4]) {
  int nx = S->nx[layer_of(S, f)];
  c[0] = f;
  c[1] = f + 1;
  c[2] = f + (uint32_t)nx;
  c[3] = f + (uint32_t)nx + 1;
}

uint32_t oc_hash_cell(int cx, int cy, int cz) {
  uint32_t h = ((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u);
  return h & OC_MASK;
}

int oc_face_valid(const oc_scene* S, uint32_t f) {
  int i, j, l;
  node_ij(S, f, &i, &j, &l);
  if (i + 1 >= S->nx[l] || j + 1 >= S->ny[l]) return 0;
  uint32_t c[4];
  face_corners(S, f, c);
  return active(S, 2 * c[0]) && active(S, 2 * c[0] + 1) && active(S, 2 * c[1] + 1) && active(S, 2 * c[2]);
}

static void centroid4(const float* a, const float* b, const float* c, const float* d, float* o) {
  for (int q = 0; q < 3; q++) o[q] = 0.25f * (((a[q] + b[q]) + c[q]) + d[q]);
}

/* the cells a face is hashed in: predicted centroid, and the start centroid when another cell */
static int face_cells(const oc_scene* S, uint32_t f, int cell[2][3]) {
  uint32_t c[4];
  face_corners(S, f, c);
  float p[4][3], p0[4][3], m1[3], m0[3];
  for (int k = 0; k < 4; k++) {
    at_now(S, c[k], p[k]);
    at_start(S, c[k], p0[k]);
  }
  centroid4(p[0], p[1], p[2], p[3], m1);
  centroid4(p0[0], p0[1], p0[2], p0[3], m0);
  for (int q = 0; q < 3; q++) {
    cell[0][q] = (int)floorf(m1[q] / OC_CELL);
    cell[1][q] = (int)floorf(m0[q] / OC_CELL);
  }
  return (cell[0][0] == cell[1][0] && cell[0][1] == cell[1][1] && cell[0][2] == cell[1][2]) ? 1 : 2;
}

/* closestTri: weights of (a, b, c) and squared distance */
float oc_closest_tri(const float* p, const float* a, const float* b, const float* c, float* w) {
  float ab[3], ac[3], ap[3];
  sub3f(b, a, ab);
  sub3f(c, a, ac);
  sub3f(p, a, ap);
  float d1 = dot3f(ab, ap), d2 = dot3f(ac, ap);
  w[0] = 1;
  w[1] = 0;
  w[2] = 0;
  if (d1 > 0 || d2 > 0) {
    float bp[3], cp[3];
    sub3f(p, b, bp);
    sub3f(p, c, cp);
    float d3 = dot3f(ab, bp), d4 = dot3f(ac, bp);
    float d5 = dot3f(ab, cp), d6 = dot3f(ac, cp);
    float vc = d1 * d4 - d3 * d2;
    float vb = d5 * d2 - d1 * d6;
    float va = d3 * d6 - d5 * d4;
    if (d3 >= 0 && d4 <= d3) {
      w[0] = 0;
      w[1] = 1;
    } else if (d6 >= 0 && d5 <= d6) {
      w[0] = 0;
      w[2] = 1;
    } else if (vc <= 0 && d1 >= 0 && d3 <= 0) {
      float v = d1 / (d1 - d3);
      w[0] = 1 - v;
      w[1] = v;
    } else if (vb <= 0 && d2 >= 0 && d6 <= 0) {
      float v = d2 / (d2 - d6);
      w[0] = 1 - v;
      w[2] = v;
    } else if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
      float v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      w[0] = 0;
      w[1] = 1 - v;
      w[2] = v;
    } else {
      float den = 1 / ((va + vb) + vc);
      float wb = vb * den, wc = vc * den;
      w[0] = (1 - wb) - wc;
      w[1] = wb;
      w[2] = wc;
    }
  }
  float e[3];
  for (int q = 0; q < 3; q++) e[q] = p[q] - ((w[0] * a[q] + w[1] * b[q]) + w[2] * c[q]);
  return dot3f(e, e);
}

/* faceHitP: corner weights (c00, c10, c01, c11), closest point, distance */
float oc_face_hit(const float* p, const float* p00, const float* p10, const float* p01, const float* p11, float* w, float* qpt) {
  float t1[3], t2[3];
  float s1 = oc_closest_tri(p, p00, p10, p11, t1);
  float s2 = oc_closest_tri(p, p00, p11, p01, t2);
  float d;
  if (s1 <= s2) {
    w[0] = t1[0];
    w[1] = t1[1];
    w[2] = 0;
    w[3] = t1[2];
    d = sqrtf(s1);
  } else {
    w[0] = t2[0];
    w[1] = 0;
    w[2] = t2[2];
    w[3] = t2[1];
    d = sqrtf(s2);
  }
  for (int q = 0; q < 3; q++) qpt[q] = ((w[0] * p00[q] + w[1] * p10[q]) + w[2] * p01[q]) + w[3] * p11[q];
  return d;
}

static float hit_face(const oc_scene* S, const float* p, uint32_t f, int start, float* w, float* qpt) {
  uint32_t c[4];
  face_corners(S, f, c);
  float P[4][3];
  for (int k = 0; k < 4; k++) {
    if (start)
      at_start(S, c[k], P[k]);
    else
      at_now(S, c[k], P[k]);
  }
  return oc_face_hit(p, P[0], P[1], P[2], P[3], w, qpt);
}

float oc_seg_dist(const float* qp, const float* a, const float* b) {
  float e[3], qa[3];
  sub3f(b, a, e);
  sub3f(qp, a, qa);
  float s = dot3f(qa, e) / fmaxf(dot3f(e, e), 1e-20f);
  s = fminf(fmaxf(s, 0), 1);
  float r[3];
  for (int q = 0; q < 3; q++) r[q] = qa[q] - s * e[q];
  return len3f(r);
}

float oc_seg_tri(const float* o, const float* d, const float* a, const float* b, const float* c) {
  float e1[3], e2[3], p[3], s[3], qv[3];
  sub3f(b, a, e1);
  sub3f(c, a, e2);
  cross3f(d, e2, p);
  float det = dot3f(e1, p);
  if (fabsf(det) < 1e-18f) return 2;
  float inv = 1 / det;
  sub3f(o, a, s);
  float u = dot3f(s, p) * inv;
  if (u < 0 || u > 1) return 2;
  cross3f(s, e1, qv);
  float v = dot3f(d, qv) * inv;
  if (v < 0 || u + v > 1) return 2;
  float t = dot3f(e2, qv) * inv;
  return (t >= 0 && t <= 1) ? t : 2;
}

static float path_cross(const oc_scene* S, const float* o, const float* d, uint32_t f) {
  uint32_t c[4];
  face_corners(S, f, c);
  float a[3], b[3], e[3], g[3];
  at_start(S, c[0], a);
  at_start(S, c[1], b);
  at_start(S, c[2], e);
  at_start(S, c[3], g);
  float t = fminf(oc_seg_tri(o, d, a, b, g), oc_seg_tri(o, d, a, g, e));
  at_now(S, c[0], a);
  at_now(S, c[1], b);
  at_now(S, c[2], e);
  at_now(S, c[3], g);
  return fminf(t, fminf(oc_seg_tri(o, d, a, b, g), oc_seg_tri(o, d, a, g, e)));
}

static uint32_t lnode(const oc_scene* S, int l, int i, int j) { return (uint32_t)(S->off[l] + j * S->nx[l] + i); }

static int row_ok(const oc_scene* S, int l, int j, int i0, int i1) {
  int lo = i0 < i1 ? i0 : i1, hi = i0 < i1 ? i1 : i0;
  for (int i = lo; i < hi; i++)
    if (!active(S, 2 * lnode(S, l, i, j))) return 0;
  return 1;
}
static int col_ok(const oc_scene* S, int l, int i, int j0, int j1) {
  int lo = j0 < j1 ? j0 : j1, hi = j0 < j1 ? j1 : j0;
  for (int j = lo; j < hi; j++)
    if (!active(S, 2 * lnode(S, l, i, j) + 1)) return 0;
  return 1;
}
static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static int linked(const oc_scene* S, int l, int ni, int nj, int fi, int fj) {
  int ci = clampi(ni, fi, fi + 1), cj = clampi(nj, fj, fj + 1);
  if (row_ok(S, l, nj, ni, ci) && col_ok(S, l, ci, nj, cj)) return 1;
  return col_ok(S, l, ni, nj, cj) && row_ok(S, l, cj, ni, ci);
}

/* damageRow: per node, a broken span within the window along its lattice row */
static void row_marks(const oc_scene* S, uint32_t* mark) {
  for (int n = 0; n < S->N; n++) {
    int i, j, l;
    node_ij(S, (uint32_t)n, &i, &j, &l);
    uint32_t m = 0;
    int a0 = i - OC_WINDOW < 0 ? 0 : i - OC_WINDOW, a1 = i + OC_WINDOW > S->nx[l] - 1 ? S->nx[l] - 1 : i + OC_WINDOW;
    for (int a = a0; a <= a1; a++) {
      uint32_t q = lnode(S, l, a, j);
      if ((S->flags[2 * q] & 2) || (S->flags[2 * q + 1] & 2)) m = 1;
    }
    mark[n] = m;
  }
}

/* the per-node query of detect over a candidate face list (the hash records, or every face) */
typedef struct {
  uint32_t best[OC_K];
  float bestD[OC_K];
  float d0min, n0[3];
  uint32_t f0;
  float tc;
  uint32_t fc;
} oc_query;

static void query_face(const oc_scene* S, oc_query* Q, const float* p, const float* p0, const float* path, int moving,
                       const float* c1, float r1, const float* c0, uint32_t f, int l, int ni, int nj, int damaged) {
  float t1[3], t0[3];
  sub3f(p, c1, t1);
  sub3f(p0, c0, t0);
  int nearNow = len3f(t1) - r1 < OC_REACH;
  int nearStart = len3f(t0) - r1 < fminf(Q->d0min, 0.5f * OC_CELL);
  int onPath = moving && fminf(oc_seg_dist(c0, p0, p), oc_seg_dist(c1, p0, p)) < r1;
  if (!nearNow && !nearStart && !onPath) return;
  int fi, fj, fl;
  node_ij(S, f, &fi, &fj, &fl);
  int di = fi - ni, dj = fj - nj;
  if (fl == l && di >= -3 && di <= 2 && dj >= -3 && dj <= 2 && (!damaged || linked(S, l, ni, nj, ni + di, nj + dj))) return;
  float w[4], qp[3];
  if (nearStart) {
    float d = hit_face(S, p0, f, 1, w, qp);
    if ((d < Q->d0min || (d == Q->d0min && f < Q->f0)) && d > 1e-9f) {
      Q->d0min = d;
      for (int q = 0; q < 3; q++) Q->n0[q] = (p0[q] - qp[q]) / d;
      Q->f0 = f;
    }
  }
  if (onPath) {
    float t = path_cross(S, p0, path, f);
    if (t < Q->tc || (t == Q->tc && t <= 1 && f < Q->fc)) {
      Q->tc = t;
      Q->fc = f;
    }
  }
  if (!nearNow) return;
  float d = hit_face(S, p, f, 0, w, qp);
  if (d >= OC_REACH) return;
  uint32_t cand = f;
  float cd = d;
  for (int k = 0; k < OC_K; k++) {
    if (cand == Q->best[k]) break;
    if (cd < Q->bestD[k] || (cd == Q->bestD[k] && cand < Q->best[k])) {
      uint32_t ts = Q->best[k];
      float td = Q->bestD[k];
      Q->best[k] = cand;
      Q->bestD[k] = cd;
      cand = ts;
      cd = td;
    }
  }
}

/* the face record (centroid now, radius, centroid at start) */
static void face_record(const oc_scene* S, uint32_t f, float* rec) {
  uint32_t c[4];
  face_corners(S, f, c);
  float a0[4][3], a1[4][3], m0[3], m1[3];
  for (int k = 0; k < 4; k++) {
    at_start(S, c[k], a0[k]);
    for (int q = 0; q < 3; q++) a1[k][q] = a0[k][q] + S->dx[3 * c[k] + q];
  }
  centroid4(a0[0], a0[1], a0[2], a0[3], m0);
  centroid4(a1[0], a1[1], a1[2], a1[3], m1);
  float r0 = 0, r1 = 0;
  for (int k = 0; k < 4; k++) {
    float t[3];
    sub3f(a0[k], m0, t);
    r0 = fmaxf(r0, dot3f(t, t));
    sub3f(a1[k], m1, t);
    r1 = fmaxf(r1, dot3f(t, t));
  }
  rec[0] = m1[0];
  rec[1] = m1[1];
  rec[2] = m1[2];
  rec[3] = sqrtf(fmaxf(r0, r1));
  rec[4] = m0[0];
  rec[5] = m0[1];
  rec[6] = m0[2];
  memcpy(&rec[7], &f, 4);
}

/* the bound of a node from its query, and its reciprocal key (0: none) */
static uint64_t finish_bound(const oc_scene* S, uint32_t n, const oc_query* Q, const float* p0, float* own, uint32_t* fb_out) {
  float nb[3] = {Q->n0[0], Q->n0[1], Q->n0[2]};
  float db = Q->d0min;
  uint32_t fb = Q->f0;
  if (Q->fc != OC_NONE && Q->fc != Q->f0) {
    uint32_t c[4];
    face_corners(S, Q->fc, c);
    float a[3], b[3], e[3], u[3], v[3], nrm[3];
    at_start(S, c[0], a);
    at_start(S, c[1], b);
    at_start(S, c[2], e);
    sub3f(b, a, u);
    sub3f(e, a, v);
    cross3f(u, v, nrm);
    float ln = len3f(nrm);
    if (ln > 1e-12f) {
      for (int q = 0; q < 3; q++) nrm[q] = nrm[q] / ln;
      float t[3];
      sub3f(p0, a, t);
      float side = dot3f(t, nrm);
      for (int q = 0; q < 3; q++) nb[q] = side >= 0 ? nrm[q] : -nrm[q];
      db = fabsf(side);
      fb = Q->fc;
    }
  }
  *fb_out = OC_NONE;
  if (fb == OC_NONE || (fb == Q->f0 && db >= 0.5f * OC_CELL)) {
    own[0] = own[1] = own[2] = 0;
    own[3] = OC_UNB;
    return 0;
  }
  float share = OC_FRACTION * fmaxf(db - OC_HARD, 0);
  own[0] = nb[0];
  own[1] = nb[1];
  own[2] = nb[2];
  own[3] = share;
  uint32_t u = (uint32_t)(share / OC_RECIP);
  *fb_out = fb;
  /* 64-bit key: the share code above a full 32-bit node + 1 (the reference packs node + 1 into 20 bits,
     selfcontact.wgsl:462-474); the same order: tightest share, then the larger node */
  return ((uint64_t)(4095u - (u < 4095u ? u : 4095u)) << 32) | (uint64_t)(n + 1u);
}

static void detect_common(const oc_scene* S, int brute, uint32_t* cells, float* items, uint32_t* pairs, float* cmeta) {
  const int N = S->N;
  uint32_t* mark = calloc((size_t)N, sizeof(uint32_t));
  uint64_t* keys = calloc((size_t)N, sizeof(uint64_t));
  uint32_t* valid = NULL;
  int nvalid = 0;
  float* recs = NULL;
  if (brute) {
    valid = malloc(sizeof(uint32_t) * (size_t)N);
    recs = malloc(sizeof(float) * 8 * (size_t)N);
    for (int f = 0; f < N; f++)
      if (oc_face_valid(S, (uint32_t)f)) {
        face_record(S, (uint32_t)f, &recs[8 * nvalid]);
        valid[nvalid++] = (uint32_t)f;
      }
  } else {
    /* faceCount, exclusive scan, faceFill (in face order) */
    memset(cells, 0, sizeof(uint32_t) * 2 * OC_TABLE);
    int cl[2][3];
    for (int f = 0; f < N; f++) {
      if (!oc_face_valid(S, (uint32_t)f)) continue;
      int k = face_cells(S, (uint32_t)f, cl);
      for (int q = 0; q < k; q++) cells[oc_hash_cell(cl[q][0], cl[q][1], cl[q][2])]++;
    }
    uint32_t run = 0;
    for (uint32_t h = 0; h < OC_TABLE; h++) {
      uint32_t c = cells[h];
      cells[h] = run;
      run += c;
    }
    for (int f = 0; f < N; f++) {
      if (!oc_face_valid(S, (uint32_t)f)) continue;
      float rec[8];
      face_record(S, (uint32_t)f, rec);
      int k = face_cells(S, (uint32_t)f, cl);
      for (int q = 0; q < k; q++) {
        uint32_t h = oc_hash_cell(cl[q][0], cl[q][1], cl[q][2]);
        uint32_t slot = cells[h] + cells[OC_TABLE + h]++;
        memcpy(&items[8 * (size_t)slot], rec, sizeof rec);
      }
    }
  }
  row_marks(S, mark);
  for (int n = 0; n < N; n++) {
    uint32_t* pr = &pairs[(OC_K + 2) * (size_t)n];
    float* own = &cmeta[8 * (size_t)n];
    for (int k = 0; k < OC_K; k++) pr[k] = OC_NONE;
    if (S->pinned[n]) {
      own[0] = own[1] = own[2] = 0;
      own[3] = OC_UNB;
      continue;
    }
    float p[3], p0[3], path[3];
    at_now(S, (uint32_t)n, p);
    at_start(S, (uint32_t)n, p0);
    sub3f(p, p0, path);
    int moving = dot3f(path, path) > OC_HARD * OC_HARD;
    oc_query Q;
    for (int k = 0; k < OC_K; k++) {
      Q.best[k] = OC_NONE;
      Q.bestD[k] = 1e9f;
    }
    Q.d0min = 1e9f;
    Q.n0[0] = Q.n0[1] = Q.n0[2] = 0;
    Q.f0 = OC_NONE;
    Q.tc = 2;
    Q.fc = OC_NONE;
    int lo[3], hi[3];
    for (int q = 0; q < 3; q++) {
      lo[q] = (int)floorf(fminf(p[q], p0[q]) / OC_CELL - 0.5f);
      long long h = (long long)(int)floorf(fmaxf(p[q], p0[q]) / OC_CELL + 0.5f);
      hi[q] = (int)(h < (long long)lo[q] + 3 ? h : (long long)lo[q] + 3);
    }
    int ni, nj, l;
    node_ij(S, (uint32_t)n, &ni, &nj, &l);
    int damaged = 0;
    int b0 = nj - OC_WINDOW < 0 ? 0 : nj - OC_WINDOW, b1 = nj + OC_WINDOW > S->ny[l] - 1 ? S->ny[l] - 1 : nj + OC_WINDOW;
    for (int b = b0; b <= b1; b++)
      if (mark[lnode(S, l, ni, b)]) damaged = 1;
    if (brute) {
      /* every valid face hashed in a cell of the scanned box (no table, no collisions) */
      int cl[2][3];
      for (int v = 0; v < nvalid; v++) {
        int k = face_cells(S, valid[v], cl), in = 0;
        for (int q = 0; q < k; q++)
          if (cl[q][0] >= lo[0] && cl[q][0] <= hi[0] && cl[q][1] >= lo[1] && cl[q][1] <= hi[1] && cl[q][2] >= lo[2] &&
              cl[q][2] <= hi[2])
            in = 1;
        if (!in) continue;
        const float* r = &recs[8 * v];
        query_face(S, &Q, p, p0, path, moving, r, r[3], r + 4, valid[v], l, ni, nj, damaged);
      }
    } else {
      for (int cz = lo[2]; cz <= hi[2]; cz++)
        for (int cy = lo[1]; cy <= hi[1]; cy++)
          for (int cx = lo[0]; cx <= hi[0]; cx++) {
            uint32_t h = oc_hash_cell(cx, cy, cz);
            uint32_t start = cells[h], count = cells[OC_TABLE + h];
            for (uint32_t r = 0; r < count; r++) {
              const float* rec = &items[8 * (size_t)(start + r)];
              uint32_t f;
              memcpy(&f, &rec[7], 4);
              query_face(S, &Q, p, p0, path, moving, rec, rec[3], rec + 4, f, l, ni, nj, damaged);
            }
          }
    }
    for (int k = 0; k < OC_K; k++) pr[k] = Q.best[k];
    uint32_t fb;
    uint64_t key = finish_bound(S, (uint32_t)n, &Q, p0, own, &fb);
    if (key) {
      uint32_t c[4];
      face_corners(S, fb, c);
      for (int k = 0; k < 4; k++)
        if (key > keys[c[k]]) keys[c[k]] = key;
    }
  }
  /* resolveBounds */
  for (int n = 0; n < N; n++) {
    float* rec = &cmeta[8 * (size_t)n + 4];
    uint64_t key = keys[n];
    pairs[(OC_K + 2) * (size_t)n + OC_K] = 0;
    pairs[(OC_K + 2) * (size_t)n + OC_K + 1] = mark[n];
    if (!key) {
      rec[0] = rec[1] = rec[2] = 0;
      rec[3] = OC_UNB;
      continue;
    }
    uint32_t j = (uint32_t)(key & 0xFFFFFFFFu) - 1;
    float share = (float)(4095u - (uint32_t)(key >> 32)) * OC_RECIP;
    for (int q = 0; q < 3; q++) rec[q] = -cmeta[8 * (size_t)j + q];
    rec[3] = share;
  }
  free(mark);
  free(keys);
  free(valid);
  free(recs);
}

/* faceCount / scan / faceFill / damageRow / detect / resolveBounds */
void oc_detect(const oc_scene* S, uint32_t* cells, float* items, uint32_t* pairs, float* cmeta) {
  detect_common(S, 0, cells, items, pairs, cmeta);
}

/* the same query over every valid face whose cell lies in the node's scan box (pins the hash) */
void oc_detect_brute(const oc_scene* S, uint32_t* pairs, float* cmeta) { detect_common(S, 1, NULL, NULL, pairs, cmeta); }

static int32_t fixq(float v, float scale) {
  float t = v * scale;
  t = fmaxf(t, -2e9f);
  t = fminf(t, 2e9f);
  return (int32_t)t;
}

static void add_term(int32_t* acc, uint32_t n, const float* g, const float* H) {
  int32_t* a = &acc[9 * (size_t)n];
  for (int q = 0; q < 3; q++) a[q] = (int32_t)((uint32_t)a[q] + (uint32_t)fixq(g[q], OC_GS));
  for (int q = 0; q < 6; q++) a[3 + q] = (int32_t)((uint32_t)a[3 + q] + (uint32_t)fixq(H[q], OC_HS));
}

/* terms + gather: penalty and friction at the current iterate into ext (acc is returned too) */
void oc_terms(const oc_scene* S, const uint32_t* pairs, float* cmeta, float h, float gnorm, int32_t* acc, float* ext) {
  const int N = S->N;
  memset(acc, 0, sizeof(int32_t) * 9 * (size_t)N);
  const float epsU = OC_MU_SPEED * h;
  static const int ri[6] = {0, 0, 0, 1, 1, 2}, ci[6] = {0, 1, 2, 1, 2, 2};
  for (int n = 0; n < N; n++) {
    float p[3];
    at_now(S, (uint32_t)n, p);
    float* own = &cmeta[8 * (size_t)n];
    float reach = 1e9f;
    float sW = sqrtf(((S->mass[n] * gnorm) * OC_DELTA) / OC_STIFF);
    const float* dn = &S->dx[3 * n];
    for (int k = 0; k < OC_K; k++) {
      uint32_t f = pairs[(OC_K + 2) * (size_t)n + k];
      if (f == OC_NONE) continue;
      uint32_t c[4];
      face_corners(S, f, c);
      float w[4], qp[3];
      float d = hit_face(S, p, f, 0, w, qp);
      if (d <= 1e-9f) continue;
      float nn[3];
      for (int q = 0; q < 3; q++) nn[q] = (p[q] - qp[q]) / d;
      float a = dot3f(nn, own);
      if (a > 0.1f) reach = fminf(reach, ((d - OC_DELTA) + fmaxf(sW, 2 * fmaxf(OC_DELTA - d, 0))) / a);
      if (d >= OC_DELTA) continue;
      float s = OC_DELTA - d;
      float lam = ((OC_STIFF * s) * s) / OC_DELTA;
      float rel[3], ut[3];
      for (int q = 0; q < 3; q++) {
        float fm = ((w[0] * S->dx[3 * c[0] + q] + w[1] * S->dx[3 * c[1] + q]) + w[2] * S->dx[3 * c[2] + q]) + w[3] * S->dx[3 * c[3] + q];
        rel[q] = dn[q] - fm;
      }
      float rn = dot3f(rel, nn);
      for (int q = 0; q < 3; q++) ut[q] = rel[q] - rn * nn[q];
      float kf = (OC_MU * lam) / fmaxf(len3f(ut), epsU);
      float g[3], H[6], H2[6];
      for (int q = 0; q < 3; q++) g[q] = (-lam) * nn[q] + kf * ut[q];
      float c1 = ((2 * OC_STIFF) * s) / OC_DELTA;
      for (int e = 0; e < 6; e++) {
        float t = nn[ri[e]] * nn[ci[e]];
        H[e] = c1 * t + kf * ((ri[e] == ci[e] ? 1.f : 0.f) - t);
        H2[e] = 2 * H[e];
      }
      add_term(acc, (uint32_t)n, g, H2);
      for (int qn = 0; qn < 4; qn++) {
        if (w[qn] <= 0) continue;
        float gq[3], Hq[6];
        for (int q = 0; q < 3; q++) gq[q] = (-w[qn]) * g[q];
        float s2 = 2 * w[qn];
        for (int e = 0; e < 6; e++) Hq[e] = s2 * H[e];
        add_term(acc, c[qn], gq, Hq);
      }
    }
    if (reach < 1e9f && (own[0] != 0 || own[1] != 0 || own[2] != 0)) {
      float approach = -dot3f(dn, own);
      own[3] = fminf(own[3], approach + reach);
    }
  }
  for (int n = 0; n < N; n++)
    for (int k = 0; k < 9; k++) ext[9 * (size_t)n + k] = (float)acc[9 * (size_t)n + k] / (k < 3 ? OC_GS : OC_HS);
}

/* bound, feasible (step.wgsl:61-77) */
static void bound1(float* d, const float* m) {
  float approach = -dot3f(d, m);
  if (approach > m[3]) {
    float e = approach - m[3];
    for (int q = 0; q < 3; q++) d[q] = d[q] + e * m[q];
  }
}
void oc_feasible(const float* cmeta_n, float* d) {
  bound1(d, cmeta_n);
  bound1(d, cmeta_n + 4);
  bound1(d, cmeta_n);
}

/* floorBound's z (step.wgsl:79-83, common.wgsl:247-265) */
float oc_floor_bound_z(float z0, float linz, float dz, float half, float mass, float gnorm, float fz) {
  float delta = half + OC_FLOOR_BULGE;
  float sW = sqrtf(((mass * gnorm) * delta) / OC_FLOOR_K);
  float zl = z0 + linz;
  float s = fmaxf(delta - (zl - fz), 0);
  float lowest = fminf(fmaxf((fz + delta) - fmaxf(sW, 2 * s), fz + half), zl);
  return fmaxf(dz, lowest - z0);
}

/* projectStart: the predicted step made feasible (dx updated in place) */
void oc_project_start(const oc_scene* S, const float* cmeta, float* dx, int floor_on, float fz, const float* half, float gnorm) {
  for (int n = 0; n < S->N; n++) {
    if (S->pinned[n]) continue;
    float d[3] = {dx[3 * n], dx[3 * n + 1], dx[3 * n + 2]};
    oc_feasible(&cmeta[8 * (size_t)n], d);
    if (floor_on) d[2] = oc_floor_bound_z(S->x[3 * n + 2], 0, d[2], half[layer_of(S, (uint32_t)n)], S->mass[n], gnorm, fz);
    memcpy(&dx[3 * n], d, sizeof d);
  }
}

/* ------------------------------------------------------------------------------------------------
 * Cross-layer contact (src/simulation/gpu/wgsl/contact.wgsl:1-164, law gpu/contact.ts:8-18, pair
 * distances gpu/contact.ts:52-59). A node hash keyed by (1 mm cell, layer) over 2^20 slots, the K
 * nearest nodes of the other layers within reach, the directional bound to the nearest foreign node,
 * and per node a quadratic penalty with regularised Coulomb friction. Ties between equally near
 * nodes go to the lower node index (the reference keeps whichever it met first; its fill order is
 * that of its atomics). dist3 holds the contact distances of layer pairs 0-1, 0-2 and 1-2. cells is
 * 2 x 2^20 (starts, then counts), items N node ids, pairs K per node, ext 9 per node.
 * ------------------------------------------------------------------------------------------------ */
#define OP_TABLE (1u << 20)
#define OP_MASK (OP_TABLE - 1u)
#define OP_CELL 1.0e-3f
#define OP_STIFF 3000.f
#define OP_MU 0.45f
#define OP_MU_SPEED 0.002f
#define OP_FRACTION 0.45f

static int op_cell_of(float p) {
  float q = floorf(p / OP_CELL);
  return (int)fminf(fmaxf(q, -1e9f), 1e9f);
}

uint32_t oc_ply_hash_cell(int cx, int cy, int cz, uint32_t layer) {
  return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u) ^ (layer * 2654435761u)) &
         OP_MASK;
}

static uint32_t op_node_hash(const oc_scene* S, uint32_t n) {
  float p[3];
  at_now(S, n, p);
  return oc_ply_hash_cell(op_cell_of(p[0]), op_cell_of(p[1]), op_cell_of(p[2]), (uint32_t)layer_of(S, n));
}

static float op_pair_distance(const float* dist3, int la, int lb) {
  int s = la + lb;
  return s == 1 ? dist3[0] : (s == 2 ? dist3[1] : dist3[2]);
}

/* hashCount, the exclusive scan and hashFill (contact.wgsl:62-81): nodes in ascending order */
static void op_build_hash(const oc_scene* S, uint32_t* cells, uint32_t* items) {
  memset(cells, 0, 2 * OP_TABLE * sizeof(uint32_t));
  for (uint32_t n = 0; n < (uint32_t)S->N; n++) cells[op_node_hash(S, n)]++;
  uint32_t run = 0;
  for (uint32_t h = 0; h < OP_TABLE; h++) {
    uint32_t c = cells[h];
    cells[h] = run;
    run += c;
  }
  for (uint32_t n = 0; n < (uint32_t)S->N; n++) {
    uint32_t h = op_node_hash(S, n);
    items[cells[h] + cells[OP_TABLE + h]++] = n;
  }
}

/* detect (contact.wgsl:83-137) */
void oc_ply_detect(const oc_scene* S, const float* dist3, uint32_t* cells, uint32_t* items, uint32_t* pairs, float* cmeta) {
  op_build_hash(S, cells, items);
  const float reach = 0.25f * OP_CELL;
  for (uint32_t n = 0; n < (uint32_t)S->N; n++) {
    float p[3], q[3], dv[3];
    at_now(S, n, p);
    int ln = layer_of(S, n);
    uint32_t best[OC_K];
    float bestD[OC_K];
    for (int k = 0; k < OC_K; k++) {
      best[k] = OC_NONE;
      bestD[k] = OC_UNB;
    }
    float dmin = OC_UNB, nmin[3] = {0.f, 0.f, 0.f};
    uint32_t mmin = OC_NONE;
    int cx = op_cell_of(p[0]), cy = op_cell_of(p[1]), cz = op_cell_of(p[2]);
    for (int lo = 0; lo < S->nl; lo++) {
      if (lo == ln) continue;
      float r = op_pair_distance(dist3, ln, lo);
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dxc = -1; dxc <= 1; dxc++) {
            uint32_t h = oc_ply_hash_cell(cx + dxc, cy + dy, cz + dz, (uint32_t)lo);
            uint32_t start = cells[h], count = cells[OP_TABLE + h];
            for (uint32_t t = 0; t < count; t++) {
              uint32_t m = items[start + t];
              if (layer_of(S, m) != lo) continue;
              at_now(S, m, q);
              sub3f(p, q, dv);
              float d = len3f(dv);
              float e = d - r;
              if ((e < dmin || (e == dmin && m < mmin)) && d > 1e-9f) {
                dmin = e;
                mmin = m;
                for (int a = 0; a < 3; a++) nmin[a] = dv[a] / d;
              }
              if (d >= r + reach) continue;
              uint32_t cand = m;
              float cd = e;
              for (int k = 0; k < OC_K; k++) {
                if (cd < bestD[k] || (cd == bestD[k] && cand < best[k])) {
                  uint32_t ts = best[k];
                  float td = bestD[k];
                  best[k] = cand;
                  bestD[k] = cd;
                  cand = ts;
                  cd = td;
                }
              }
            }
          }
    }
    for (int k = 0; k < OC_K; k++) pairs[OC_K * n + k] = best[k];
    float* own = &cmeta[8 * n];
    if (dmin < 1e8f) {
      own[0] = nmin[0];
      own[1] = nmin[1];
      own[2] = nmin[2];
      own[3] = OP_FRACTION * (fmaxf(0.f, dmin) + 0.5f * dist3[0]);
    } else {
      own[0] = own[1] = own[2] = 0.f;
      own[3] = OC_UNB;
    }
    cmeta[8 * n + 4] = cmeta[8 * n + 5] = cmeta[8 * n + 6] = 0.f;
    cmeta[8 * n + 7] = OC_UNB;
  }
}

/* terms (contact.wgsl:139-164): ext = gradient xyz, Hessian xx xy xz yy yz zz per node */
void oc_ply_terms(const oc_scene* S, const float* dist3, const uint32_t* pairs, float h, float* ext) {
  const float eps_u = OP_MU_SPEED * h;
  for (uint32_t n = 0; n < (uint32_t)S->N; n++) {
    float g[3] = {0.f, 0.f, 0.f}, H[3][3] = {{0.f}};
    float p[3], q[3], dv[3];
    at_now(S, n, p);
    int ln = layer_of(S, n);
    for (int k = 0; k < OC_K; k++) {
      uint32_t m = pairs[OC_K * n + k];
      if (m == OC_NONE) continue;
      at_now(S, m, q);
      sub3f(p, q, dv);
      float d = len3f(dv);
      float r = op_pair_distance(dist3, ln, layer_of(S, m));
      if (d >= r || d <= 1e-9f) continue;
      float nn[3] = {dv[0] / d, dv[1] / d, dv[2] / d};
      float lam = OP_STIFF * (r - d);
      for (int a = 0; a < 3; a++) g[a] = g[a] - lam * nn[a];
      float nnT[3][3];
      for (int c = 0; c < 3; c++)
        for (int rr = 0; rr < 3; rr++) nnT[c][rr] = nn[rr] * nn[c];
      for (int c = 0; c < 3; c++)
        for (int rr = 0; rr < 3; rr++) H[c][rr] = H[c][rr] + OP_STIFF * nnT[c][rr];
      float rel[3], ut[3];
      for (int a = 0; a < 3; a++) rel[a] = S->dx[3 * n + a] - S->dx[3 * m + a];
      float rn = (rel[0] * nn[0] + rel[1] * nn[1]) + rel[2] * nn[2];
      for (int a = 0; a < 3; a++) ut[a] = rel[a] - rn * nn[a];
      float kf = OP_MU * lam / fmaxf(len3f(ut), eps_u);
      for (int a = 0; a < 3; a++) g[a] = g[a] + kf * ut[a];
      for (int c = 0; c < 3; c++)
        for (int rr = 0; rr < 3; rr++) H[c][rr] = H[c][rr] + kf * ((rr == c ? 1.f : 0.f) - nnT[c][rr]);
    }
    float* e = &ext[9 * n];
    e[0] = g[0];
    e[1] = g[1];
    e[2] = g[2];
    e[3] = H[0][0];
    e[4] = H[0][1];
    e[5] = H[0][2];
    e[6] = H[1][1];
    e[7] = H[1][2];
    e[8] = H[2][2];
  }
}

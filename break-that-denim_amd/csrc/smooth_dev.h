// Device code of the line smoother shared by its kernels (smooth.hip) and the one-workgroup solver
// (wgsolve.hip): the pass arguments, the branch-free row gather, the block parallel cyclic
// reduction of one line and the line geometry. The algorithm is described in smooth.hip.
#pragma once
#include "bd_internal.h"
#include "oprd.h"

namespace bd {

namespace {

struct SmoothArgs {
  const float* op;       // in the pass's layout (column-major for v-lines of a dual-layout level)
  const uint8_t* fixed;  // same layout
  int Np;
  const int4* lines;     // (start, stride, length, family), start in the pass's layout
  float* x;              // same layout
  const float* b;        // same layout
  float* xo;             // the other layout's copy of x (dual-layout levels), or null
  const int* gate;
  Geo geo;
  // the iterate the gather reads (same layout as x): x itself when lines of one colour share no
  // coupling (3 colours); on a one-colour level the other buffer of a pair, so that every line reads
  // the lagged iterate, not the one its neighbours are writing
  const float* xr = nullptr;
  // two line colours: lines two apart share a colour, so their couplings (the +-2 across the line)
  // are majorised onto the diagonal like the +-2 along it (newtonMultigrid.ts:704-705)
  int maj_across = 0;
};

// index strides of lattice offsets (di, dj) in the pass's layout
template <int FAM, bool DUAL>
__device__ __forceinline__ int stride_i(int nx, int ny) { return (FAM == 1 && DUAL) ? ny : 1; }
template <int FAM, bool DUAL>
__device__ __forceinline__ int stride_j(int nx, int ny) { return (FAM == 1 && DUAL) ? 1 : nx; }

__device__ __forceinline__ void sym_full(const float* s, float* o) {
  o[0] = s[0]; o[1] = s[1]; o[2] = s[2];
  o[3] = s[1]; o[4] = s[3]; o[5] = s[4];
  o[6] = s[2]; o[7] = s[4]; o[8] = s[5];
}

// inverse of the symmetric block shifted by PIVOT_SHIFT * max|diag| (newtonMultigrid.ts:414-423)
__device__ __forceinline__ void pivot_inv(const float* s, float* o) {
  float e = (float)kPivotShift * fmaxf(fabsf(s[0]), fmaxf(fabsf(s[3]), fabsf(s[5])));
  float a00 = s[0] + e, a01 = s[1], a02 = s[2], a11 = s[3] + e, a12 = s[4], a22 = s[5] + e;
  float c00 = a11 * a22 - a12 * a12, c01 = a12 * a02 - a01 * a22, c02 = a01 * a12 - a11 * a02;
  // hardware reciprocal (1 ulp): the pivots of a preconditioner need no correctly rounded division
  float id = __builtin_amdgcn_rcpf(a00 * c00 + a01 * c01 + a02 * c02);
  o[0] = c00 * id;
  o[1] = c01 * id;
  o[2] = c02 * id;
  o[3] = (a00 * a22 - a02 * a02) * id;
  o[4] = (a02 * a01 - a00 * a12) * id;
  o[5] = (a00 * a11 - a01 * a01) * id;
}

__device__ __forceinline__ void mm(const float* a, const float* b, float* o) {
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) o[r * 3 + c] = a[r * 3] * b[c] + a[r * 3 + 1] * b[3 + c] + a[r * 3 + 2] * b[6 + c];
}
__device__ __forceinline__ void mtm(const float* a, const float* b, float* o) {  // a^T b
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) o[r * 3 + c] = a[r] * b[c] + a[3 + r] * b[3 + c] + a[6 + r] * b[6 + c];
}
__device__ __forceinline__ void mmt(const float* a, const float* b, float* o) {  // a b^T
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) o[r * 3 + c] = a[r * 3] * b[c * 3] + a[r * 3 + 1] * b[c * 3 + 1] + a[r * 3 + 2] * b[c * 3 + 2];
}
__device__ __forceinline__ void mv(const float* a, const float* v, float* o) {
  o[0] = a[0] * v[0] + a[1] * v[1] + a[2] * v[2];
  o[1] = a[3] * v[0] + a[4] * v[1] + a[5] * v[2];
  o[2] = a[6] * v[0] + a[7] * v[1] + a[8] * v[2];
}
__device__ __forceinline__ void mtv(const float* a, const float* v, float* o) {
  o[0] = a[0] * v[0] + a[3] * v[1] + a[6] * v[2];
  o[1] = a[1] * v[0] + a[4] * v[1] + a[7] * v[2];
  o[2] = a[2] * v[0] + a[5] * v[1] + a[8] * v[2];
}
__device__ __forceinline__ void smv(const float* s, const float* v, float* o) {
  o[0] = s[0] * v[0] + s[1] * v[1] + s[2] * v[2];
  o[1] = s[1] * v[0] + s[3] * v[1] + s[4] * v[2];
  o[2] = s[2] * v[0] + s[4] * v[1] + s[5] * v[2];
}

// one coupling of the row gather: source row and plane base of its symmetric block (or scalar)
struct Cpl {
  int di, dj;   // offset of the neighbour
  int sdi, sdj; // offset of the row that stores the block (0,0: own forward planes)
  int base;     // first plane
  bool scalar;
};

// fine stencil: 4 forward blocks, their 4 mirrors, 2 forward scalars and their mirrors
__device__ constexpr Cpl kFine[12] = {
    {1, 0, 0, 0, F_E, false},   {-1, 0, -1, 0, F_E, false}, {0, 1, 0, 0, F_N, false},    {0, -1, 0, -1, F_N, false},
    {1, 1, 0, 0, F_NE, false},  {-1, -1, -1, -1, F_NE, false}, {-1, 1, 0, 0, F_NW, false}, {1, -1, 1, -1, F_NW, false},
    {2, 0, 0, 0, F_EE, true},   {-2, 0, -2, 0, F_EE, true}, {0, 2, 0, 0, F_NN, true},    {0, -2, 0, -2, F_NN, true}};

// What gather_row computes: the whole row, only its right-hand side r (which changes with the
// iterate every pass), or only its line blocks D and U (which do not: the operator's own rows).
constexpr int kGatherAll = 0, kGatherR = 1, kGatherDU = 2;

// Gathers D (6), U (9: +1 along coupling, full), r (3) of row n on a line of family FAM (PART: all of
// them, r only, or D and U only; the parts it computes are the same bits in every mode).
// All loads use clamped in-bounds indices and are masked afterwards, so they issue together.
template <bool FINE, int FAM, bool DUAL, class RD, int PART = kGatherAll>
__device__ __forceinline__ void gather_row(const SmoothArgs& S, const RD& rd, int n, int nb, int i, int j, int nx, int ny,
                                           float D[6], float U[9], float r[3], bool& fixed_n) {
  const int Np = S.Np;
  const float* x = S.xr;
  const int si = stride_i<FAM, DUAL>(nx, ny), sj = stride_j<FAM, DUAL>(nx, ny);
  fixed_n = S.fixed[n] != 0;
  float c[6];
  if (PART != kGatherR) rd.blk(F_C, n, c);
  float xn0 = 0.f, xn1 = 0.f, xn2 = 0.f;
  if (PART != kGatherDU) {
    xn0 = x[n];
    xn1 = x[Np + n];
    xn2 = x[2 * Np + n];
    r[0] = S.b[n];
    r[1] = S.b[Np + n];
    r[2] = S.b[2 * Np + n];
  }
#pragma unroll
  for (int q = 0; q < 9; q++) U[q] = 0.f;
  float maj = 0.f;
  auto one = [&](int di, int dj, int sdi, int sdj, int base, bool scalar, bool neg_src) {
    const bool in = i + di >= 0 && i + di < nx && j + dj >= 0 && j + dj < ny;
    const int m = in ? n + dj * sj + di * si : n;
    const int src = in ? n + sdj * sj + sdi * si : n;
    float blk[6];
    if (scalar) {
      float s = rd.scalar(base, src);
      blk[0] = s; blk[1] = 0.f; blk[2] = 0.f; blk[3] = s; blk[4] = 0.f; blk[5] = s;
    } else {
      rd.blk(base, src, blk);
    }
    const float xm[3] = {x[m], x[Np + m], x[2 * Np + m]};
    // the fixed-neighbour mask is arithmetic: a branch on the loaded flag would split the gather
    // into one memory round trip per coupling
    const uint8_t fm = S.fixed[m];  // m is clamped in bounds: the load is unconditional
    const float w = (in & (fm == 0)) ? 1.f : 0.f;
    const int along = FAM == 0 ? di : dj;
    const int across = FAM == 0 ? dj : di;
    if (across == 0 && along == 1) {
      if (PART != kGatherR) {
        float full[9];
        sym_full(blk, full);
#pragma unroll
        for (int q = 0; q < 9; q++) U[q] = w * full[q];
      }
    } else if (!(across == 0 && along == -1)) {
      if (PART != kGatherDU) {
        float t[3];
        smv(blk, xm, t);
        r[0] -= w * t[0];
        r[1] -= w * t[1];
        r[2] -= w * t[2];
      }
      // +-2 along the line; with two colours also +-2 across it
      if (across == 0 || (S.maj_across && (across == 2 || across == -2))) maj += w * sym_frob(blk);
    }
  };
  if (FINE) {
#pragma unroll
    for (int k = 0; k < 12; k++) one(kFine[k].di, kFine[k].dj, kFine[k].sdi, kFine[k].sdj, kFine[k].base, kFine[k].scalar, false);
  } else {
    // 24 couplings of 6-plane blocks: batches of 6 keep the loads in flight within the VGPR budget
#pragma unroll 3
    for (int k = 1; k < kCoarseBlocks; k++) {
      one(coarse_di(k), coarse_dj(k), 0, 0, 6 * k, false, false);
      one(-coarse_di(k), -coarse_dj(k), -coarse_di(k), -coarse_dj(k), 6 * k, false, true);
    }
  }
  if (fixed_n) {
    if (PART != kGatherR) {
      D[0] = 1.f; D[1] = 0.f; D[2] = 0.f; D[3] = 1.f; D[4] = 0.f; D[5] = 1.f;
#pragma unroll
      for (int q = 0; q < 9; q++) U[q] = 0.f;
    }
    if (PART != kGatherDU) {
      r[0] = S.b[n];
      r[1] = S.b[Np + n];
      r[2] = S.b[2 * Np + n];
    }
    return;
  }
  if (PART != kGatherR) {
#pragma unroll
    for (int q = 0; q < 6; q++) D[q] = c[q];
    D[0] += maj;
    D[3] += maj;
    D[5] += maj;
  }
  if (PART != kGatherDU) {
    r[0] += maj * xn0;
    r[1] += maj * xn1;
    r[2] += maj * xn2;
  }
}

// Partial row gather of rows on one wave's 64 lanes for the couplings c with c % G == GI (GI is a
// compile-time constant, so the part is straight-line code whose loads all issue together). GI 0
// also loads the row's centre block, b, x and fixed flag. Writes the partial rhs correction and
// majorisation to LDS; the part owning the +1 along-line coupling writes U. Lanes past the line's
// end (valid false) gather a clamped row and write nothing.
template <bool FINE, int FAM, int G, int GI, bool DUAL, class RD>
__device__ __forceinline__ void gather_part(const SmoothArgs& S, const RD& rd, int n, int nb, int i, int j, int nx, int ny,
                                            bool valid, int row, int len, float* pr, float* pm, float* gU, float* gD,
                                            float* gR, float* gX, uint8_t* gF) {
  const int Np = S.Np;
  const float* x = S.xr;
  const int si = stride_i<FAM, DUAL>(nx, ny), sj = stride_j<FAM, DUAL>(nx, ny);
  float r0 = 0.f, r1 = 0.f, r2 = 0.f, maj = 0.f;
  float U[9];
  bool ownsU = false;
  auto one = [&](int di, int dj, int sdi, int sdj, int base, bool scalar) {
    const bool in = i + di >= 0 && i + di < nx && j + dj >= 0 && j + dj < ny;
    const int m = in ? n + dj * sj + di * si : n;
    const int src = in ? n + sdj * sj + sdi * si : n;
    float blk[6];
    if (scalar) {
      float s = rd.scalar(base, src);
      blk[0] = s; blk[1] = 0.f; blk[2] = 0.f; blk[3] = s; blk[4] = 0.f; blk[5] = s;
    } else {
      rd.blk(base, src, blk);
    }
    const float xm[3] = {x[m], x[Np + m], x[2 * Np + m]};
    const uint8_t fm = S.fixed[m];  // m is clamped in bounds: the load is unconditional
    const float w = (in & (fm == 0)) ? 1.f : 0.f;
    const int along = FAM == 0 ? di : dj;
    const int across = FAM == 0 ? dj : di;
    if (across == 0 && along == 1) {
      float full[9];
      sym_full(blk, full);
#pragma unroll
      for (int q = 0; q < 9; q++) U[q] = w * full[q];
      ownsU = true;
    } else if (!(across == 0 && along == -1)) {
      float t[3];
      smv(blk, xm, t);
      r0 += w * t[0];
      r1 += w * t[1];
      r2 += w * t[2];
      if (across == 0 || (S.maj_across && (across == 2 || across == -2))) maj += w * sym_frob(blk);
    }
  };
  if (FINE) {
#pragma unroll
    for (int k = 0; k < 12; k++)
      if (k % G == GI) one(kFine[k].di, kFine[k].dj, kFine[k].sdi, kFine[k].sdj, kFine[k].base, kFine[k].scalar);
  } else {
#pragma unroll
    for (int k = 0; k < 24; k++) {
      if (k % G != GI) continue;
      const int b = 1 + (k >> 1);
      if (k & 1)
        one(-coarse_di(b), -coarse_dj(b), -coarse_di(b), -coarse_dj(b), 6 * b, false);
      else
        one(coarse_di(b), coarse_dj(b), 0, 0, 6 * b, false);
    }
  }
  float c[6], bb[3], xx[3];
  uint8_t f = 0;
  if (GI == 0) {
    rd.blk(F_C, n, c);
#pragma unroll
    for (int q = 0; q < 3; q++) {
      bb[q] = S.b[q * Np + n];
      xx[q] = x[q * Np + n];
    }
    f = S.fixed[n];
  }
  if (!valid) return;
  pr[(0 * G + GI) * len + row] = r0;
  pr[(1 * G + GI) * len + row] = r1;
  pr[(2 * G + GI) * len + row] = r2;
  pm[GI * len + row] = maj;
  if (ownsU) {
#pragma unroll
    for (int q = 0; q < 9; q++) gU[q * len + row] = U[q];
  }
  if (GI == 0) {
#pragma unroll
    for (int q = 0; q < 6; q++) gD[q * len + row] = c[q];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      gR[q * len + row] = bb[q];
      gX[q * len + row] = xx[q];
    }
    gF[row] = f;
  }
}

// Runs part g (wave-uniform) of the gather.
template <bool FINE, int FAM, int G, bool DUAL, class RD, int GI = 0>
__device__ __forceinline__ void gather_dispatch(int g, const SmoothArgs& S, const RD& rd, int n, int nb, int i, int j, int nx,
                                                int ny, bool valid, int row, int len, float* pr, float* pm, float* gU,
                                                float* gD, float* gR, float* gX, uint8_t* gF) {
  if constexpr (GI < G) {
    if (g == GI)
      gather_part<FINE, FAM, G, GI, DUAL>(S, rd, n, nb, i, j, nx, ny, valid, row, len, pr, pm, gU, gD, gR, gX, gF);
    else
      gather_dispatch<FINE, FAM, G, DUAL, RD, GI + 1>(g, S, rd, n, nb, i, j, nx, ny, valid, row, len, pr, pm, gU, gD, gR,
                                                      gX, gF);
  }
}

// LDS record of one reduced row in the line solve: D^-1 (6, symmetric), U (9), r (3) and 2 floats of
// padding, so that a record is five 16-byte vector accesses with immediate offsets
constexpr int kRowF = 20;
// LDS floats per reduced row: two record buffers
constexpr int kSolveFloats = 2 * kRowF;
// variant codes of the split smoother (row kernel + solve kernel, two or four rows per thread; the
// factored one solves with the line's elimination kept from the operator's factorization)
constexpr int kSplitVariant = 16;
constexpr int kSplit4Variant = 17;
constexpr int kFactorVariant = 18;

__device__ __forceinline__ void put_row(float* B, int t, const float Di[6], const float U[9], const float r[3]) {
  float4* p = reinterpret_cast<float4*>(B + kRowF * t);
  p[0] = make_float4(Di[0], Di[1], Di[2], Di[3]);
  p[1] = make_float4(Di[4], Di[5], U[0], U[1]);
  p[2] = make_float4(U[2], U[3], U[4], U[5]);
  p[3] = make_float4(U[6], U[7], U[8], r[0]);
  p[4] = make_float4(r[1], r[2], 0.f, 0.f);
}

__device__ __forceinline__ void get_row(const float* B, int j, float Di[6], float U[9], float r[3]) {
  const float4* p = reinterpret_cast<const float4*>(B + kRowF * j);
  const float4 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4];
  Di[0] = a.x; Di[1] = a.y; Di[2] = a.z; Di[3] = a.w; Di[4] = b.x; Di[5] = b.y;
  U[0] = b.z; U[1] = b.w; U[2] = c.x; U[3] = c.y; U[4] = c.z; U[5] = c.w; U[6] = d.x; U[7] = d.y; U[8] = d.z;
  r[0] = d.w; r[1] = e.x; r[2] = e.y;
}

// Block parallel cyclic reduction of R reduced rows of a line, one per thread (t < R): each row's
// symmetric D, its forward coupling U (to row t + 1) and r in registers. At level d every row
// eliminates both its neighbours at distance d at once, so after ceil(log2 R) levels (one barrier
// each, the row kept in registers, LDS ping-ponged between the record buffers B0 and B1) each row
// stands alone; there is no back-substitution sweep. The reduced system stays symmetric (row k's
// coupling to k - d is row k - d's coupling to k, transposed), so a row carries only its forward
// coupling, and each row publishes its own shifted pivot inverse with it (one inversion per row and
// level). The first level writes B0 and the next one B1, so one barrier per level orders every read
// before the overwrite; a caller may read B1 before the call. Returns the row's solution x and the
// next row's xn (zero past the last row).
__device__ __forceinline__ void pcr_rows(float* B0, float* B1, int t, int R, float D[6], float U[9], float r[3],
                                         float x[3], float xn[3]) {
  const bool row = t < R;
  int lvl = 0;
  for (int d = 1; d < R; d <<= 1, lvl ^= 1) {
    float* B = lvl ? B1 : B0;
    if (row) {
      float Di[6];
      pivot_inv(D, Di);
      put_row(B, t, Di, U, r);
    }
    __syncthreads();
    if (row) {
      float Un[9];
#pragma unroll
      for (int q = 0; q < 9; q++) Un[q] = 0.f;
      const int j1 = t - d, j2 = t + d;
      if (j1 >= 0) {
        // row j1 couples to this row by U1: D -= U1^T D1^-1 U1, r -= U1^T D1^-1 r1
        float Di[6], U1[9], r1[3], T[9];
        get_row(B, j1, Di, U1, r1);
        float Dif[9];
        sym_full(Di, Dif);
        mm(Dif, U1, T);  // T = D1^-1 U1
        D[0] -= U1[0] * T[0] + U1[3] * T[3] + U1[6] * T[6];
        D[1] -= U1[0] * T[1] + U1[3] * T[4] + U1[6] * T[7];
        D[2] -= U1[0] * T[2] + U1[3] * T[5] + U1[6] * T[8];
        D[3] -= U1[1] * T[1] + U1[4] * T[4] + U1[7] * T[7];
        D[4] -= U1[1] * T[2] + U1[4] * T[5] + U1[7] * T[8];
        D[5] -= U1[2] * T[2] + U1[5] * T[5] + U1[8] * T[8];
        float tt[3];
        mtv(T, r1, tt);  // U1^T D1^-1 r1
        r[0] -= tt[0];
        r[1] -= tt[1];
        r[2] -= tt[2];
      }
      if (j2 < R) {
        // this row couples to row j2 by U, row j2 on to j2 + d by U2: D -= U D2^-1 U^T,
        // r -= U D2^-1 r2, and the new coupling to j2 + d is -U D2^-1 U2
        float Di[6], U2[9], r2[3], T2[9];
        get_row(B, j2, Di, U2, r2);
        float Dif[9];
        sym_full(Di, Dif);
        mm(U, Dif, T2);  // T2 = U D2^-1
        D[0] -= T2[0] * U[0] + T2[1] * U[1] + T2[2] * U[2];
        D[1] -= T2[0] * U[3] + T2[1] * U[4] + T2[2] * U[5];
        D[2] -= T2[0] * U[6] + T2[1] * U[7] + T2[2] * U[8];
        D[3] -= T2[3] * U[3] + T2[4] * U[4] + T2[5] * U[5];
        D[4] -= T2[3] * U[6] + T2[4] * U[7] + T2[5] * U[8];
        D[5] -= T2[6] * U[6] + T2[7] * U[7] + T2[8] * U[8];
        float tt[3];
        mv(T2, r2, tt);
        r[0] -= tt[0];
        r[1] -= tt[1];
        r[2] -= tt[2];
        float W[9];
        mm(T2, U2, W);
#pragma unroll
        for (int q = 0; q < 9; q++) Un[q] = -W[q];
      }
#pragma unroll
      for (int q = 0; q < 9; q++) U[q] = Un[q];
    }
  }
  // every reduced row now stands alone
  float* B = lvl ? B1 : B0;
  x[0] = x[1] = x[2] = 0.f;
  if (row) {
    float inv[6];
    pivot_inv(D, inv);
    smv(inv, r, x);
    B[kRowF * t] = x[0];
    B[kRowF * t + 1] = x[1];
    B[kRowF * t + 2] = x[2];
  }
  __syncthreads();
  xn[0] = xn[1] = xn[2] = 0.f;
  if (t + 1 < R) {
    xn[0] = B[kRowF * (t + 1)];
    xn[1] = B[kRowF * (t + 1) + 1];
    xn[2] = B[kRowF * (t + 1) + 2];
  }
}

// Block tridiagonal solve of one line, two rows per thread.
// Thread t holds rows a = 2t and b = 2t+1 (D sym, U full coupling to the next row, r) and
// receives their solutions. The first odd-even level eliminates row b in registers (neighbours
// exchange one 9-float term through LDS), leaving R2 = ceil(len/2) reduced rows, one per thread,
// for pcr_rows; then b is recovered from a's and the next thread's solutions.
// cr: kSolveFloats * Lr floats of LDS.
__device__ __forceinline__ void line_solve(float* cr, int Lr, int t, int len, const float* Da_in, const float* Ua,
                                           const float* ra_in, const float* Db, const float* Ub, const float* rb,
                                           float xa[3], float xb[3]) {
  float* B0 = cr;
  float* B1 = cr + kRowF * Lr;  // also carries the register level's exchange (its first 9 floats)
  const int R2 = (len + 1) >> 1;
  const bool has_a = 2 * t < len, has_b = 2 * t + 1 < len;
  float D[6], r[3], U[9];
#pragma unroll
  for (int q = 0; q < 6; q++) D[q] = Da_in[q];
#pragma unroll
  for (int q = 0; q < 3; q++) r[q] = ra_in[q];

  // register level: eliminate row b (U_a couples a -> b, U_b couples b -> 2t+2)
  float Dib[6], Dibf[9], yb[3];
  pivot_inv(Db, Dib);
  sym_full(Dib, Dibf);
  smv(Dib, rb, yb);
  {
    float T[9], P[9], p[3];
    mm(Dibf, Ub, T);
    mtm(Ub, T, P);
    mtv(Ub, yb, p);
    if (has_b && t + 1 < R2) {
      float4* e = reinterpret_cast<float4*>(B1 + kRowF * t);
      e[0] = make_float4(P[0], 0.5f * (P[1] + P[3]), 0.5f * (P[2] + P[6]), P[4]);
      e[1] = make_float4(0.5f * (P[5] + P[7]), P[8], p[0], p[1]);
      e[2] = make_float4(p[2], 0.f, 0.f, 0.f);
    }
  }
  {
    float T2[9], W[9], t3[3];
    mm(Ua, Dibf, T2);
    mmt(T2, Ua, W);
    D[0] -= W[0];
    D[1] -= 0.5f * (W[1] + W[3]);
    D[2] -= 0.5f * (W[2] + W[6]);
    D[3] -= W[4];
    D[4] -= 0.5f * (W[5] + W[7]);
    D[5] -= W[8];
    mv(T2, rb, t3);
    r[0] -= t3[0];
    r[1] -= t3[1];
    r[2] -= t3[2];
    mm(T2, Ub, W);
#pragma unroll
    for (int q = 0; q < 9; q++) U[q] = -W[q];
  }
  __syncthreads();
  if (has_a && t >= 1) {
    const float4* e = reinterpret_cast<const float4*>(B1 + kRowF * (t - 1));
    const float4 e0 = e[0], e1 = e[1], e2 = e[2];
    D[0] -= e0.x; D[1] -= e0.y; D[2] -= e0.z; D[3] -= e0.w; D[4] -= e1.x; D[5] -= e1.y;
    r[0] -= e1.z; r[1] -= e1.w; r[2] -= e2.x;
  }
  // B1's exchange terms are read before the first level's barrier (inside pcr_rows)
  float xn[3];
  pcr_rows(B0, B1, t, R2, D, U, r, xa, xn);
  float tt[3], q1[3];
  mtv(Ua, xa, tt);
  mv(Ub, xn, q1);
  float rr[3] = {rb[0] - tt[0] - q1[0], rb[1] - tt[1] - q1[1], rb[2] - tt[2] - q1[2]};
  smv(Dib, rr, xb);
}

// Eliminates row b (D_b, r_b; coupled to row a by U_a, to the next row by U_b) from the pair: row a
// gains D_a -= U_a D_b^-1 U_a^T, r_a -= U_a D_b^-1 r_b and the coupling Un = -U_a D_b^-1 U_b to the
// next row; the next row gains P = U_b^T D_b^-1 U_b (symmetric, 6) and p = U_b^T D_b^-1 r_b, returned
// for the caller to subtract. Dib: D_b's shifted pivot inverse, kept for recovering x_b.
__device__ __forceinline__ void eliminate_pair(float Da[6], float ra[3], const float Ua[9], const float Db[6],
                                               const float Ub[9], const float rb[3], float Dib[6], float Un[9],
                                               float P[6], float p[3]) {
  float Dibf[9], yb[3], T[9], W[9], t3[3];
  pivot_inv(Db, Dib);
  sym_full(Dib, Dibf);
  smv(Dib, rb, yb);
  mm(Dibf, Ub, T);
  mtm(Ub, T, W);
  P[0] = W[0];
  P[1] = 0.5f * (W[1] + W[3]);
  P[2] = 0.5f * (W[2] + W[6]);
  P[3] = W[4];
  P[4] = 0.5f * (W[5] + W[7]);
  P[5] = W[8];
  mtv(Ub, yb, p);
  float T2[9];
  mm(Ua, Dibf, T2);
  mmt(T2, Ua, W);
  Da[0] -= W[0];
  Da[1] -= 0.5f * (W[1] + W[3]);
  Da[2] -= 0.5f * (W[2] + W[6]);
  Da[3] -= W[4];
  Da[4] -= 0.5f * (W[5] + W[7]);
  Da[5] -= W[8];
  mv(T2, rb, t3);
  ra[0] -= t3[0];
  ra[1] -= t3[1];
  ra[2] -= t3[2];
  mm(T2, Ub, W);
#pragma unroll
  for (int q = 0; q < 9; q++) Un[q] = -W[q];
}

// x_b = D_b^-1 (r_b - U_a^T x_a - U_b x_n): a row eliminated by eliminate_pair, recovered from the
// solutions of its neighbours.
__device__ __forceinline__ void recover_row(const float Dib[6], const float Ua[9], const float Ub[9], const float rb[3],
                                            const float xa[3], const float xn[3], float xb[3]) {
  float tt[3], q1[3];
  mtv(Ua, xa, tt);
  mv(Ub, xn, q1);
  const float rr[3] = {rb[0] - tt[0] - q1[0], rb[1] - tt[1] - q1[1], rb[2] - tt[2] - q1[2]};
  smv(Dib, rr, xb);
}

// Block tridiagonal solve of one line, four rows per thread: thread t holds rows 4t + m (m = 0..3,
// a b c d; D sym, U full coupling to the next row, r; rows past the line are identity) and receives
// their solutions. Two odd-even levels run in registers: b is eliminated into a and c, d into c and
// the next thread's a, then c into a and the next thread's a. What a thread adds to the next
// thread's a depends on its own rows only, so both levels' terms go through one LDS exchange. That
// leaves R4 = ceil(len/4) reduced rows, one per thread, for pcr_rows: a quarter of the line's rows in
// the parallel reduction, whose work grows as rows x levels, against half for line_solve. Then c, b
// and d are recovered from a's and the next thread's solutions. With pivots shifted at each
// elimination, the result equals line_solve's up to the shifts' different order and rounding.
// cr: kSolveFloats * Lr floats of LDS.
__device__ __forceinline__ void line_solve4(float* cr, int Lr, int t, int len, float (&D)[4][6], float (&U)[4][9],
                                            float (&r)[4][3], float (&x)[4][3]) {
  float* B0 = cr;
  float* B1 = cr + kRowF * Lr;  // also carries the register levels' exchange (its first 9 floats)
  const int R4 = (len + 3) >> 2;
  // level 1: b into a (and c), d into c (and the next thread's a)
  float Dib[6], Uac[9], Pb[6], pb[3];
  eliminate_pair(D[0], r[0], U[0], D[1], U[1], r[1], Dib, Uac, Pb, pb);
  float Did[6], Uce[9], Pd[6], pd[3];
  eliminate_pair(D[2], r[2], U[2], D[3], U[3], r[3], Did, Uce, Pd, pd);
#pragma unroll
  for (int q = 0; q < 6; q++) D[2][q] -= Pb[q];
#pragma unroll
  for (int q = 0; q < 3; q++) r[2][q] -= pb[q];
  // level 2: c into a (and the next thread's a); a then couples to the next thread's a by Uae
  const float rc[3] = {r[2][0], r[2][1], r[2][2]};
  float Dic[6], Uae[9], Pc[6], pc[3];
  eliminate_pair(D[0], r[0], Uac, D[2], Uce, rc, Dic, Uae, Pc, pc);
  if (4 * t < len && t + 1 < R4) {
    float4* e = reinterpret_cast<float4*>(B1 + kRowF * t);
    e[0] = make_float4(Pd[0] + Pc[0], Pd[1] + Pc[1], Pd[2] + Pc[2], Pd[3] + Pc[3]);
    e[1] = make_float4(Pd[4] + Pc[4], Pd[5] + Pc[5], pd[0] + pc[0], pd[1] + pc[1]);
    e[2] = make_float4(pd[2] + pc[2], 0.f, 0.f, 0.f);
  }
  __syncthreads();
  if (t >= 1 && t < R4) {
    const float4* e = reinterpret_cast<const float4*>(B1 + kRowF * (t - 1));
    const float4 e0 = e[0], e1 = e[1], e2 = e[2];
    D[0][0] -= e0.x; D[0][1] -= e0.y; D[0][2] -= e0.z; D[0][3] -= e0.w; D[0][4] -= e1.x; D[0][5] -= e1.y;
    r[0][0] -= e1.z; r[0][1] -= e1.w; r[0][2] -= e2.x;
  }
  // B1's exchange terms are read before the first level's barrier (inside pcr_rows)
  float xe[3];
  pcr_rows(B0, B1, t, R4, D[0], Uae, r[0], x[0], xe);
  recover_row(Dic, Uac, Uce, rc, x[0], xe, x[2]);
  recover_row(Dib, U[0], U[1], r[1], x[0], x[2], x[1]);
  recover_row(Did, U[2], U[3], r[3], x[2], xe, x[3]);
}

// ---- factored line solve ----
// A line's block-tridiagonal matrix does not change between the passes over one operator, only r
// does. So the elimination (four rows per thread: two odd-even levels in registers, then cyclic
// reduction over the threads' first rows) runs once per operator and keeps, per row, what a pass
// needs at the point the row is eliminated: its shifted pivot inverse Di (6), G = U_l Di (9; U_l
// the coupling from the row's left neighbour at that point) and H = Di U_r (9; U_r its coupling to
// the right neighbour). A pass then moves r only: eliminating row e subtracts G r_e from its left
// neighbour's r and H^T r_e from its right one's; its solution is x_e = Di r_e - G^T x_l - H x_r.
// kFacPlanes planes per row: Di, G, H.
constexpr int kFacPlanes = 24;

__device__ __forceinline__ void sym_part(const float W[9], float P[6]) {
  P[0] = W[0];
  P[1] = 0.5f * (W[1] + W[3]);
  P[2] = 0.5f * (W[2] + W[6]);
  P[3] = W[4];
  P[4] = 0.5f * (W[5] + W[7]);
  P[5] = W[8];
}

// Eliminates a row (D, coupled from its left neighbour by Ul and to its right neighbour by Ur): its
// factor (Di, G = Ul Di, H = Di Ur), the left neighbour's update Wl = Ul Di Ul^T, the right one's
// Pr = Ur^T Di Ur (both symmetric, to subtract) and the new coupling Un = -Ul Di Ur between them.
__device__ __forceinline__ void factor_row(const float D[6], const float Ul[9], const float Ur[9], float Di[6],
                                           float G[9], float H[9], float Wl[6], float Pr[6], float Un[9]) {
  float Dif[9], W[9];
  pivot_inv(D, Di);
  sym_full(Di, Dif);
  mm(Ul, Dif, G);
  mm(Dif, Ur, H);
  mmt(G, Ul, W);
  sym_part(W, Wl);
  mtm(Ur, H, W);
  sym_part(W, Pr);
  mm(Ul, H, W);
#pragma unroll
  for (int q = 0; q < 9; q++) Un[q] = -W[q];
}

// the row's factor to its planes (rows past the line are not written)
__device__ __forceinline__ void put_factor(float* of, size_t cap, int k, int len, const float Di[6], const float G[9],
                                           const float H[9]) {
  if (k >= len) return;
#pragma unroll
  for (int q = 0; q < 6; q++) of[q * cap + k] = Di[q];
#pragma unroll
  for (int q = 0; q < 9; q++) of[(6 + q) * cap + k] = G[q];
#pragma unroll
  for (int q = 0; q < 9; q++) of[(15 + q) * cap + k] = H[q];
}

__device__ __forceinline__ void sub6(float D[6], const float P[6]) {
#pragma unroll
  for (int q = 0; q < 6; q++) D[q] -= P[q];
}

// Factorization of one line (thread t: rows 4t + m, D sym and U to the next row from the line blocks;
// rows past the line are identity rows): writes every row's factor. ex: 8 floats of LDS per thread.
// Every thread of the block calls it.
__device__ __forceinline__ void line_factor4(float* ex, float* of, size_t cap, int t, int len, float (&D)[4][6],
                                             const float (&U)[4][9]) {
  const int k0 = 4 * t;
  const int R4 = (len + 3) >> 2;
  float Di[6], G[9], H[9], Wl[6], Pr[6];
  // level 1: b between a and c, d between c and the next thread's a
  float Uac[9], Uce[9], Pd[6];
  factor_row(D[1], U[0], U[1], Di, G, H, Wl, Pr, Uac);
  put_factor(of, cap, k0 + 1, len, Di, G, H);
  sub6(D[0], Wl);
  sub6(D[2], Pr);
  factor_row(D[3], U[2], U[3], Di, G, H, Wl, Pd, Uce);
  put_factor(of, cap, k0 + 3, len, Di, G, H);
  sub6(D[2], Wl);
  // level 2: c between a and the next thread's a
  float Pc[6], Ua[9];
  factor_row(D[2], Uac, Uce, Di, G, H, Wl, Pc, Ua);
  put_factor(of, cap, k0 + 2, len, Di, G, H);
  sub6(D[0], Wl);
  float* e = ex + 8 * t;
  if (t < R4) {
#pragma unroll
    for (int q = 0; q < 6; q++) e[q] = Pd[q] + Pc[q];
  }
  __syncthreads();
  if (t >= 1 && t < R4) {
    const float* p = ex + 8 * (t - 1);
#pragma unroll
    for (int q = 0; q < 6; q++) D[0][q] -= p[q];
  }
  // cyclic reduction of the R4 first rows: at stride s the rows t = s mod 2s are eliminated between
  // t - s and t + s. An eliminated row publishes Di (6), H (9) and Pr = U^T Di U (6); its left
  // neighbour forms and writes its G. Two barriers a level (the buffer is reused).
  for (int s = 1; s < R4; s <<= 1) {
    __syncthreads();
    const int m = t & (2 * s - 1);
    if (t < R4 && m == s) {
      float Dif[9], W[9], P[6];
      pivot_inv(D[0], Di);
      sym_full(Di, Dif);
      mm(Dif, Ua, H);
      mtm(Ua, H, W);
      sym_part(W, P);
#pragma unroll
      for (int q = 0; q < 6; q++) of[q * cap + k0] = Di[q];
#pragma unroll
      for (int q = 0; q < 9; q++) of[(15 + q) * cap + k0] = H[q];
      float* w = ex + 24 * t;
#pragma unroll
      for (int q = 0; q < 6; q++) w[q] = Di[q];
#pragma unroll
      for (int q = 0; q < 9; q++) w[6 + q] = H[q];
#pragma unroll
      for (int q = 0; q < 6; q++) w[15 + q] = P[q];
    }
    __syncthreads();
    if (t < R4 && m == 0) {
      if (t + s < R4) {
        const float* w = ex + 24 * (t + s);
        float Dif[9], W[9], Hr[9];
        sym_full(w, Dif);
#pragma unroll
        for (int q = 0; q < 9; q++) Hr[q] = w[6 + q];
        mm(Ua, Dif, G);
        mmt(G, Ua, W);
        sym_part(W, Wl);
        sub6(D[0], Wl);
        const int kr = 4 * (t + s);
#pragma unroll
        for (int q = 0; q < 9; q++) of[(6 + q) * cap + kr] = G[q];
        mm(Ua, Hr, W);
#pragma unroll
        for (int q = 0; q < 9; q++) Ua[q] = -W[q];
      }
      if (t >= s) sub6(D[0], ex + 24 * (t - s) + 15);
    }
  }
  // the last row standing
  if (t == 0 && len > 0) {
    pivot_inv(D[0], Di);
#pragma unroll
    for (int q = 0; q < 9; q++) G[q] = H[q] = 0.f;
    put_factor(of, cap, 0, len, Di, G, H);
  }
}

// One pass of a factored line (thread t: rows 4t + m): F the rows' factors (Di 0-5, G 6-14, H 15-23;
// zero past the line), r their right-hand sides (zero past the line); returns x. lds: 6 float4 per
// thread of the block. Every thread of the block calls it.
__device__ __forceinline__ void line_solve_factored(float4* lds, int Lr, int t, int len, const float (&F)[kFacPlanes][4],
                                                    float (&r)[4][3], float (&x)[4][3]) {
  const int R4 = (len + 3) >> 2;
  float4* Q = lds;
  float4* CL = lds + Lr;      // two buffers: contributions to the left neighbour
  float4* CRt = lds + 3 * Lr; // two buffers: contributions to the right neighbour
  float4* X = lds + 5 * Lr;
  float Fm[4][kFacPlanes];
#pragma unroll
  for (int m = 0; m < 4; m++)
#pragma unroll
    for (int q = 0; q < kFacPlanes; q++) Fm[m][q] = F[q][m];
  x[0][0] = x[0][1] = x[0][2] = 0.f;
  float v[3];
  // level 1: b into a and c; d into c and the next thread's a
  mv(Fm[1] + 6, r[1], v);
  r[0][0] -= v[0]; r[0][1] -= v[1]; r[0][2] -= v[2];
  mtv(Fm[1] + 15, r[1], v);
  r[2][0] -= v[0]; r[2][1] -= v[1]; r[2][2] -= v[2];
  mv(Fm[3] + 6, r[3], v);
  r[2][0] -= v[0]; r[2][1] -= v[1]; r[2][2] -= v[2];
  float q3[3], v2[3];
  mtv(Fm[3] + 15, r[3], q3);
  // level 2: c into a and the next thread's a
  mv(Fm[2] + 6, r[2], v);
  r[0][0] -= v[0]; r[0][1] -= v[1]; r[0][2] -= v[2];
  mtv(Fm[2] + 15, r[2], v2);
  if (t < R4) Q[t] = make_float4(q3[0] + v2[0], q3[1] + v2[1], q3[2] + v2[2], 0.f);
  __syncthreads();
  if (t >= 1 && t < R4) {
    const float4 p = Q[t - 1];
    r[0][0] -= p.x; r[0][1] -= p.y; r[0][2] -= p.z;
  }
  // cyclic reduction of the first rows, one barrier a level (buffers alternate)
  int lvl = 0;
  int top = 0;
  for (int s = 1; s < R4; s <<= 1, lvl ^= 1) {
    top = s;
    const int m = t & (2 * s - 1);
    if (t < R4 && m == s) {
      float a[3], b[3];
      mv(Fm[0] + 6, r[0], a);
      mtv(Fm[0] + 15, r[0], b);
      CL[lvl * Lr + t] = make_float4(a[0], a[1], a[2], 0.f);
      CRt[lvl * Lr + t] = make_float4(b[0], b[1], b[2], 0.f);
    }
    __syncthreads();
    if (t < R4 && m == 0) {
      if (t + s < R4) {
        const float4 p = CL[lvl * Lr + t + s];
        r[0][0] -= p.x; r[0][1] -= p.y; r[0][2] -= p.z;
      }
      if (t >= s) {
        const float4 p = CRt[lvl * Lr + t - s];
        r[0][0] -= p.x; r[0][1] -= p.y; r[0][2] -= p.z;
      }
    }
  }
  // back substitution: the last row standing, then the rows of each level, widest stride first
  if (t == 0 && R4 > 0) {
    smv(Fm[0], r[0], x[0]);
    X[0] = make_float4(x[0][0], x[0][1], x[0][2], 0.f);
  }
  __syncthreads();
  for (int s = top; s >= 1; s >>= 1) {
    const int m = t & (2 * s - 1);
    if (t < R4 && m == s) {
      const float4 xl = X[t - s];
      const float4 xr = t + s < R4 ? X[t + s] : make_float4(0.f, 0.f, 0.f, 0.f);
      const float l3[3] = {xl.x, xl.y, xl.z}, r3[3] = {xr.x, xr.y, xr.z};
      float a[3], b[3];
      smv(Fm[0], r[0], x[0]);
      mtv(Fm[0] + 6, l3, a);
      mv(Fm[0] + 15, r3, b);
      x[0][0] -= a[0] + b[0]; x[0][1] -= a[1] + b[1]; x[0][2] -= a[2] + b[2];
      X[t] = make_float4(x[0][0], x[0][1], x[0][2], 0.f);
    }
    __syncthreads();
  }
  float xe[3] = {0.f, 0.f, 0.f};
  if (t + 1 < R4) {
    const float4 p = X[t + 1];
    xe[0] = p.x; xe[1] = p.y; xe[2] = p.z;
  }
  // the register levels' rows: c from a and e, then b from a and c, d from c and e
  float a[3], b[3];
  smv(Fm[2], r[2], x[2]);
  mtv(Fm[2] + 6, x[0], a);
  mv(Fm[2] + 15, xe, b);
  x[2][0] -= a[0] + b[0]; x[2][1] -= a[1] + b[1]; x[2][2] -= a[2] + b[2];
  smv(Fm[1], r[1], x[1]);
  mtv(Fm[1] + 6, x[0], a);
  mv(Fm[1] + 15, x[2], b);
  x[1][0] -= a[0] + b[0]; x[1][1] -= a[1] + b[1]; x[1][2] -= a[2] + b[2];
  smv(Fm[3], r[3], x[3]);
  mtv(Fm[3] + 6, x[2], a);
  mv(Fm[3] + 15, xe, b);
  x[3][0] -= a[0] + b[0]; x[3][1] -= a[1] + b[1]; x[3][2] -= a[2] + b[2];
}

__device__ __forceinline__ void set_identity(float D[6], float U[9], float r[3]) {
  D[0] = 1.f; D[1] = 0.f; D[2] = 0.f; D[3] = 1.f; D[4] = 0.f; D[5] = 1.f;
#pragma unroll
  for (int q = 0; q < 9; q++) U[q] = 0.f;
  r[0] = r[1] = r[2] = 0.f;
}

struct LineGeo {
  int4 ln;
  int nx, ny, off, i0, j0;
};

template <int FAM, bool DUAL>
__device__ __forceinline__ LineGeo line_geo_of(const SmoothArgs& S, int4 ln) {
  LineGeo g;
  g.ln = ln;
  int l = geo_layer(S.geo, g.ln.x);
  g.nx = S.geo.nx[l];
  g.ny = S.geo.ny[l];
  g.off = S.geo.off[l];
  int r0 = g.ln.x - g.off;
  if (FAM == 1 && DUAL) {
    g.i0 = r0 / g.ny;
    g.j0 = r0 - g.i0 * g.ny;
  } else {
    g.j0 = r0 / g.nx;
    g.i0 = r0 - g.j0 * g.nx;
  }
  return g;
}

template <int FAM, bool DUAL>
__device__ __forceinline__ LineGeo line_geo_at(const SmoothArgs& S, int idx) {
  return line_geo_of<FAM, DUAL>(S, S.lines[idx]);
}

// A pass's convergence gate and its line's table entry, loaded together at the start of a kernel (the
// kernels test both before their first use, so neither load waits for the other's round trip).
struct PassHead {
  int4 ln;
  bool gated;
};
__device__ __forceinline__ PassHead pass_head(const SmoothArgs& S, int idx) {
  PassHead h;
  h.ln = S.lines[idx];
  h.gated = S.gate ? *S.gate != 0 : false;
  return h;
}

// row k of the line: its index in the pass's layout, in the row-major layout (b), and in the other
// layout (the second copy of x on dual-layout levels)
template <int FAM, bool DUAL>
__device__ __forceinline__ void row_index(const LineGeo& L, int k, int& n, int& nb, int& no, int& i, int& j) {
  n = L.ln.x + k * L.ln.y;
  i = FAM == 0 ? L.i0 + k : L.i0;
  j = FAM == 0 ? L.j0 : L.j0 + k;
  const int rm = L.off + j * L.nx + i, cm = L.off + i * L.ny + j;
  nb = (FAM == 1 && DUAL) ? rm : n;
  no = DUAL ? (FAM == 1 ? rm : cm) : n;
}

// rows k0 .. k0 + 3 of the line (clamped to its last row): their indices in the pass's layout and in
// the other one, and their fixed flags, loaded before the solve so the stores need not wait for them
template <int FAM, bool DUAL>
__device__ __forceinline__ void line_rows_fixed(const SmoothArgs& S, const LineGeo& L, int k0, int len, int (&nn)[4],
                                                int (&no)[4], uint8_t (&fx)[4]) {
#pragma unroll
  for (int m = 0; m < 4; m++) {
    int nb, i, j;
    row_index<FAM, DUAL>(L, min(k0 + m, len - 1), nn[m], nb, no[m], i, j);
    fx[m] = S.fixed[nn[m]];
  }
}

__device__ __forceinline__ void store_x(const SmoothArgs& S, int n, int no, const float v[3], bool dual) {
  const int Np = S.Np;
  S.x[n] = v[0];
  S.x[Np + n] = v[1];
  S.x[2 * Np + n] = v[2];
  if (dual && S.xo) {
    S.xo[no] = v[0];
    S.xo[Np + no] = v[1];
    S.xo[2 * Np + no] = v[2];
  }
}

}  // namespace

}  // namespace bd

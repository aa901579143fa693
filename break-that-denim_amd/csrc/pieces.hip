// Detached pieces: their rigid modes in the Newton solve, and the node state.
//
// A piece (nodes connected through active spans that no mount or grip anchor reaches) keeps its
// rigid motion resisted only by its inertia, drag and contact, against span couplings ~1e5 times
// stiffer. The multigrid hierarchy does not carry such near-null modes, and in FP32 the operator
// cannot resolve them: two PCG iterations moved a falling piece by a few percent of its gravity step
// and torn pieces hovered. Restated from src/simulation/gpu/wgsl:
//   k_rigid_frame    rigid_frame.wgsl:52-161  per piece, its mass centroid at the iterate and
//                    E = Z^T K Z over its six rigid modes Z = [I, -[r]x] (K: each node's external
//                    block and its own terms' rotation stiffness, written by the assembly), then E^-1
//   k_rigid_reduce   rigid_apply.wgsl:40-80   per piece, force and torque of the rhs about the
//                    centroid and the rigid step (t, w) = E^-1 (force, torque)
//   k_rigid_apply    rigid_apply.wgsl:82-91   the PCG start x0 = t + w x r on the piece's nodes (the
//                    residual b - A x0 follows in FP64 with the PCG's true-residual kernel)
//   k_spin_reduce    rigid_spin.wgsl:51-124   per piece, centroid, momentum and inertia tensor at
//                    the substep start, and the rotation over the substep w h = I^-1 L h
//   k_spin_apply     rigid_spin.wgsl:126-139  the Newton start moved onto that rotation:
//                    dx += 2 (1 - cos th) k x (k x r)
//   k_curve_set      rigid_spin.wgsl:141-155  per node, its piece's rotation in the direction and
//                    its arm from the centroid, for the line search's curved steps
//   k_curve_apply    energy.wgsl:293-302      delta += rotationCurve(alpha w, r) / alpha
//   k_apply_state    state.wgsl:16-35         the retired bit of `pinned` from the node state
// Sums run in FP64 (the reference's are FP32 tree sums); E is factored in FP64 with the reference's
// regularisation of the rotation block.
#include <stdexcept>

#include "bd_internal.h"

namespace bd {

namespace {

constexpr int kWG = 256;  // threads of a per-piece workgroup (4 waves)

struct PieceArgs {
  int N, Np, count;
  const uint32_t* start;
  const uint32_t* nodes;
  const uint32_t* of;
  const float* x;
  float* dx;
  const float* v;
  const float* mass;
  const uint8_t* state;
  const uint8_t* pinned;
  const float* rigidk;
  double* frame;
  double* step;
  float* curve;
  double* delta;
  const float* r;
  const Scal* scal;
  double h;
};

__device__ __forceinline__ void ld3(const float* p, int Np, uint32_t n, double o[3]) {
  o[0] = p[n];
  o[1] = p[Np + n];
  o[2] = p[2 * Np + n];
}
__device__ __forceinline__ void pos(const PieceArgs& A, uint32_t n, double o[3]) {
  const int Np = A.Np;
  o[0] = (double)A.x[n] + (double)A.dx[n];
  o[1] = (double)A.x[Np + n] + (double)A.dx[Np + n];
  o[2] = (double)A.x[2 * Np + n] + (double)A.dx[2 * Np + n];
}
__device__ __forceinline__ void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ bool in_scene(const PieceArgs& A, uint32_t n) { return A.state[n] < kStateRetired; }

// sum of NV doubles over the workgroup; every thread gets the totals
template <int NV>
__device__ void wg_sum(double (&v)[NV], double (&sh)[NV][kWG / 64]) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double s = v[q];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) sh[q][wid] = s;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double s = 0;
    for (int w = 0; w < kWG / 64; w++) s += sh[q][w];
    v[q] = s;
  }
  __syncthreads();
}

// upper-triangle index of (i, j), i <= j, in a symmetric 6 x 6
__device__ __forceinline__ int tri(int i, int j) { return i * 6 - (i * (i + 1)) / 2 + j; }

// symmetric block from 6 planes (xx xy xz yy yz zz) times a vector
__device__ __forceinline__ void sym_mulv(const double s[6], const double v[3], double o[3]) {
  o[0] = s[0] * v[0] + s[1] * v[1] + s[2] * v[2];
  o[1] = s[1] * v[0] + s[3] * v[1] + s[4] * v[2];
  o[2] = s[2] * v[0] + s[4] * v[1] + s[5] * v[2];
}

__global__ __launch_bounds__(kWG) void k_rigid_frame(PieceArgs A) {
  __shared__ double sh4[4][kWG / 64];
  __shared__ double sh21[21][kWG / 64];
  const int Np = A.Np;
  for (int p = blockIdx.x; p < A.count; p += gridDim.x) {
    const uint32_t a = A.start[p], b = A.start[p + 1];
    double m[4] = {0, 0, 0, 0};
    for (uint32_t k = a + threadIdx.x; k < b; k += kWG) {
      const uint32_t n = A.nodes[k];
      if (!in_scene(A, n)) continue;
      double q[3], w = A.mass[n];
      pos(A, n, q);
      m[0] += w * q[0];
      m[1] += w * q[1];
      m[2] += w * q[2];
      m[3] += w;
    }
    wg_sum<4>(m, sh4);
    double c[3] = {0, 0, 0};
    if (m[3] > 0)
      for (int q = 0; q < 3; q++) c[q] = m[q] / m[3];
    // E = sum Z^T K Z with Z = [I, -[r]x] (a rigid displacement t + w x r), plus the rotation blocks
    double e[21];
    for (int q = 0; q < 21; q++) e[q] = 0;
    for (uint32_t k = a + threadIdx.x; k < b; k += kWG) {
      const uint32_t n = A.nodes[k];
      if (!in_scene(A, n)) continue;
      double K[6], Kr[6], r[3], Z[6][3];
      for (int q = 0; q < 6; q++) {
        K[q] = A.rigidk[q * Np + n];
        Kr[q] = A.rigidk[(6 + q) * Np + n];
      }
      pos(A, n, r);
      for (int q = 0; q < 3; q++) r[q] -= c[q];
      for (int i = 0; i < 3; i++) {
        double ei[3] = {i == 0 ? 1.0 : 0.0, i == 1 ? 1.0 : 0.0, i == 2 ? 1.0 : 0.0};
        for (int q = 0; q < 3; q++) Z[i][q] = ei[q];
        cross3(ei, r, Z[3 + i]);
      }
      for (int i = 0; i < 6; i++) {
        double KZ[3];
        sym_mulv(K, Z[i], KZ);
        for (int j = i; j < 6; j++) e[tri(i, j)] += Z[j][0] * KZ[0] + Z[j][1] * KZ[1] + Z[j][2] * KZ[2];
      }
      const int sidx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
      for (int i = 0; i < 3; i++)
        for (int j = i; j < 3; j++) e[tri(3 + i, 3 + j)] += Kr[sidx[i][j]];
    }
    wg_sum<21>(e, sh21);
    if (threadIdx.x == 0) {
      double* F = A.frame + (size_t)p * kPieceFrame;
      F[0] = c[0];
      F[1] = c[1];
      F[2] = c[2];
      F[3] = m[3] > 0 ? 1.0 : 0.0;
      // Cholesky of E, the rotation block regularised relative to its scale (a node, or a straight
      // strand, has no stiffness about its own axis, and no torque reaches that mode either)
      const double st = (e[tri(0, 0)] + e[tri(1, 1)] + e[tri(2, 2)]) / 3.0;
      const double sr = (e[tri(3, 3)] + e[tri(4, 4)] + e[tri(5, 5)]) / 3.0;
      const double eps = 1e-6 * fmax(sr, 1e-8 * st) + 1e-30;
      double M[6][6], Lc[6][6];
      for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
          M[i][j] = e[tri(i < j ? i : j, i < j ? j : i)];
          Lc[i][j] = 0;
        }
      for (int i = 3; i < 6; i++) M[i][i] += eps;
      bool ok = st > 0;
      for (int i = 0; i < 6; i++)
        for (int j = 0; j <= i; j++) {
          double s = M[i][j];
          for (int k = 0; k < j; k++) s -= Lc[i][k] * Lc[j][k];
          if (i == j) {
            if (!(s > 0)) {
              ok = false;
              s = 1;
            }
            Lc[i][i] = sqrt(s);
          } else {
            Lc[i][j] = s / Lc[j][j];
          }
        }
      for (int col = 0; col < 6; col++) {
        double y[6], z[6];
        for (int i = 0; i < 6; i++) {
          double s = i == col ? 1.0 : 0.0;
          for (int k = 0; k < i; k++) s -= Lc[i][k] * y[k];
          y[i] = s / Lc[i][i];
        }
        for (int i = 5; i >= 0; i--) {
          double s = y[i];
          for (int k = i + 1; k < 6; k++) s -= Lc[k][i] * z[k];
          z[i] = s / Lc[i][i];
        }
        for (int i = 0; i < 6; i++) F[4 + 6 * i + col] = ok ? z[i] : 0.0;
      }
    }
    __syncthreads();
  }
}

// per piece: (t, w) = E^-1 (sum r, sum (pos - c) x r) over its nodes in the scene
__global__ __launch_bounds__(kWG) void k_rigid_reduce(PieceArgs A) {
  __shared__ double sh6[6][kWG / 64];
  const int Np = A.Np;
  for (int p = blockIdx.x; p < A.count; p += gridDim.x) {
    const uint32_t a = A.start[p], b = A.start[p + 1];
    const double* F = A.frame + (size_t)p * kPieceFrame;
    double g[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t k = a + threadIdx.x; k < b; k += kWG) {
      const uint32_t n = A.nodes[k];
      if (!in_scene(A, n)) continue;
      double rn[3], q[3], t[3];
      ld3(A.r, Np, n, rn);
      pos(A, n, q);
      for (int z = 0; z < 3; z++) q[z] -= F[z];
      cross3(q, rn, t);
      for (int z = 0; z < 3; z++) {
        g[z] += rn[z];
        g[3 + z] += t[z];
      }
    }
    wg_sum<6>(g, sh6);
    if (threadIdx.x == 0) {
      double* S = A.step + (size_t)p * 8;
      const bool valid = F[3] > 0.5;
      for (int i = 0; i < 6; i++) {
        double s = 0;
        for (int j = 0; j < 6; j++) s += F[4 + 6 * i + j] * g[j];
        S[i < 3 ? i : i + 1] = valid ? s : 0.0;
      }
      S[3] = 0;
      S[7] = 0;
    }
    __syncthreads();
  }
}

__device__ __forceinline__ bool piece_node(const PieceArgs& A, uint32_t n, uint32_t& p) {
  p = A.of[n];
  return p != kNoPiece && p < (uint32_t)A.count && in_scene(A, n);
}

__global__ __launch_bounds__(256) void k_rigid_apply(PieceArgs A) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t p;
  if (n >= (uint32_t)A.N || !piece_node(A, n, p)) return;
  const double* F = A.frame + (size_t)p * kPieceFrame;
  const double* S = A.step + (size_t)p * 8;
  double q[3], w[3] = {S[4], S[5], S[6]}, t[3];
  pos(A, n, q);
  for (int z = 0; z < 3; z++) q[z] -= F[z];
  cross3(w, q, t);
  const int Np = A.Np;
  for (int z = 0; z < 3; z++) A.delta[z * Np + n] = S[z] + t[z];
}

// per piece at the substep start: centroid, momentum, inertia tensor and th = I^-1 L h
__global__ __launch_bounds__(kWG) void k_spin_reduce(PieceArgs A) {
  __shared__ double sh7[7][kWG / 64];
  __shared__ double sh9[9][kWG / 64];
  const int Np = A.Np;
  for (int p = blockIdx.x; p < A.count; p += gridDim.x) {
    const uint32_t a = A.start[p], b = A.start[p + 1];
    double s[7] = {0, 0, 0, 0, 0, 0, 0};  // m x (3), m v (3), m
    for (uint32_t k = a + threadIdx.x; k < b; k += kWG) {
      const uint32_t n = A.nodes[k];
      if (!in_scene(A, n)) continue;
      double m = A.mass[n], xq[3], vq[3];
      ld3(A.x, Np, n, xq);
      ld3(A.v, Np, n, vq);
      for (int z = 0; z < 3; z++) {
        s[z] += m * xq[z];
        s[3 + z] += m * vq[z];
      }
      s[6] += m;
    }
    wg_sum<7>(s, sh7);
    const double M = s[6];
    double c[3] = {0, 0, 0}, vc[3] = {0, 0, 0};
    if (M > 0)
      for (int z = 0; z < 3; z++) {
        c[z] = s[z] / M;
        vc[z] = s[3 + z] / M;
      }
    double t[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // L (3), I xx yy zz xy xz yz
    for (uint32_t k = a + threadIdx.x; k < b; k += kWG) {
      const uint32_t n = A.nodes[k];
      if (!in_scene(A, n)) continue;
      double m = A.mass[n], r[3], u[3], L[3];
      ld3(A.x, Np, n, r);
      ld3(A.v, Np, n, u);
      for (int z = 0; z < 3; z++) {
        r[z] -= c[z];
        u[z] -= vc[z];
      }
      cross3(r, u, L);
      for (int z = 0; z < 3; z++) t[z] += m * L[z];
      const double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      t[3] += m * (rr - r[0] * r[0]);
      t[4] += m * (rr - r[1] * r[1]);
      t[5] += m * (rr - r[2] * r[2]);
      t[6] -= m * r[0] * r[1];
      t[7] -= m * r[0] * r[2];
      t[8] -= m * r[1] * r[2];
    }
    wg_sum<9>(t, sh9);
    if (threadIdx.x == 0) {
      // regularised as the rigid frame: a node, or a straight strand about its own axis, has no
      // inertia there (and no momentum about it)
      const double eps = 1e-6 * (t[3] + t[4] + t[5]) / 3.0 + 1e-30;
      const double a00 = t[3] + eps, a11 = t[4] + eps, a22 = t[5] + eps, a01 = t[6], a02 = t[7], a12 = t[8];
      const double c00 = a11 * a22 - a12 * a12, c01 = a02 * a12 - a01 * a22, c02 = a01 * a12 - a02 * a11;
      const double c11 = a00 * a22 - a02 * a02, c12 = a01 * a02 - a00 * a12, c22 = a00 * a11 - a01 * a01;
      const double det = a00 * c00 + a01 * c01 + a02 * c02;
      double th[3] = {(c00 * t[0] + c01 * t[1] + c02 * t[2]) / det * A.h, (c01 * t[0] + c11 * t[1] + c12 * t[2]) / det * A.h,
                      (c02 * t[0] + c12 * t[1] + c22 * t[2]) / det * A.h};
      if (!(M > 0) || !isfinite(th[0]) || !isfinite(th[1]) || !isfinite(th[2])) th[0] = th[1] = th[2] = 0;
      double* S = A.step + (size_t)p * 8;
      S[0] = c[0];
      S[1] = c[1];
      S[2] = c[2];
      S[3] = 0;
      S[4] = th[0];
      S[5] = th[1];
      S[6] = th[2];
      S[7] = 0;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_spin_apply(PieceArgs A) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t p;
  if (n >= (uint32_t)A.N || !piece_node(A, n, p)) return;
  const double* S = A.step + (size_t)p * 8;
  const double th[3] = {S[4], S[5], S[6]};
  const double phi = sqrt(th[0] * th[0] + th[1] * th[1] + th[2] * th[2]);
  if (phi < 1e-9) return;
  const double k[3] = {th[0] / phi, th[1] / phi, th[2] / phi};
  double r[3], kr[3], kkr[3];
  ld3(A.x, A.Np, n, r);
  for (int z = 0; z < 3; z++) r[z] -= S[z];
  cross3(k, r, kr);
  cross3(k, kr, kkr);  // = -r_perp
  const double s = 2.0 * (1.0 - cos(phi));
  const int Np = A.Np;
  for (int z = 0; z < 3; z++) A.dx[z * Np + n] = (float)((double)A.dx[z * Np + n] + s * kkr[z]);
}

__global__ __launch_bounds__(256) void k_curve_set(PieceArgs A) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N) return;
  float w[4] = {0.f, 0.f, 0.f, 0.f}, r[3] = {0.f, 0.f, 0.f};
  uint32_t p;
  if (piece_node(A, n, p)) {
    const double* F = A.frame + (size_t)p * kPieceFrame;
    if (F[3] > 0.5) {
      const double* S = A.step + (size_t)p * 8;
      double q[3];
      pos(A, n, q);
      for (int z = 0; z < 3; z++) {
        w[z] = (float)S[4 + z];
        r[z] = (float)(q[z] - F[z]);
      }
      w[3] = 1.f;
    }
  }
  const int Np = A.Np;
  for (int z = 0; z < 4; z++) A.curve[z * Np + n] = w[z];
  for (int z = 0; z < 3; z++) A.curve[(4 + z) * Np + n] = r[z];
}

__global__ __launch_bounds__(256) void k_curve_apply(PieceArgs A) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (uint32_t)A.N || A.pinned[n]) return;
  const int Np = A.Np;
  if (A.curve[3 * Np + n] < 0.5f) return;
  const double alpha = A.scal->line_alpha;
  if (!(alpha > 0)) return;
  double th[3], r[3], c[3];
  for (int z = 0; z < 3; z++) {
    th[z] = alpha * (double)A.curve[z * Np + n];
    r[z] = A.curve[(4 + z) * Np + n];
  }
  rotation_curve(th, r, c);
  for (int z = 0; z < 3; z++) A.delta[z * Np + n] += c[z] / alpha;
}

__global__ __launch_bounds__(256) void k_apply_state(int N, const uint8_t* state, uint8_t* pinned) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  pinned[n] = (uint8_t)((pinned[n] & kPinMount) | (state[n] == kStateRetired ? kPinRetired : 0));
}

PieceArgs make_args(Solver& s) {
  PieceArgs A;
  A.N = s.N;
  A.Np = s.Np;
  A.count = s.npieces;
  A.start = s.piece_start;
  A.nodes = s.piece_nodes;
  A.of = s.piece_of;
  A.x = s.x;
  A.dx = s.dx;
  A.v = s.v;
  A.mass = s.mass;
  A.state = s.nstate;
  A.pinned = s.pinned;
  A.rigidk = s.rigidk;
  A.frame = s.piece_frame;
  A.step = s.piece_step;
  A.curve = s.curve;
  A.delta = s.delta;
  A.r = s.lv[0].b;
  A.scal = s.scal;
  A.h = s.P.h;
  return A;
}

int piece_grid(int count) { return count < 1024 ? (count > 0 ? count : 1) : 1024; }

}  // namespace

void pieces_alloc(Solver& s) {
  const size_t Np = (size_t)s.Np;
  BD_HIP(hipMalloc(&s.piece_start, (kMaxPieces + 1) * sizeof(uint32_t)));
  BD_HIP(hipMalloc(&s.piece_nodes, ((size_t)s.N + 1) * sizeof(uint32_t)));
  BD_HIP(hipMalloc(&s.piece_of, Np * sizeof(uint32_t)));
  BD_HIP(hipMalloc(&s.piece_frame, (size_t)kMaxPieces * kPieceFrame * sizeof(double)));
  BD_HIP(hipMalloc(&s.piece_step, (size_t)kMaxPieces * 8 * sizeof(double)));
  BD_HIP(hipMalloc(&s.rigidk, 12 * Np * sizeof(float)));
  BD_HIP(hipMalloc(&s.curve, 7 * Np * sizeof(float)));
  BD_HIP(hipMemsetAsync(s.piece_start, 0, (kMaxPieces + 1) * sizeof(uint32_t), s.stream));
  BD_HIP(hipMemsetAsync(s.piece_of, 0xFF, Np * sizeof(uint32_t), s.stream));
  BD_HIP(hipMemsetAsync(s.piece_frame, 0, (size_t)kMaxPieces * kPieceFrame * sizeof(double), s.stream));
  BD_HIP(hipMemsetAsync(s.piece_step, 0, (size_t)kMaxPieces * 8 * sizeof(double), s.stream));
  BD_HIP(hipMemsetAsync(s.rigidk, 0, 12 * Np * sizeof(float), s.stream));
  BD_HIP(hipMemsetAsync(s.curve, 0, 7 * Np * sizeof(float), s.stream));
}

void launch_apply_state(Solver& s) {
  k_apply_state<<<grid_for(s.N, 256), 256, 0, s.stream>>>(s.N, s.nstate, s.pinned);
}

void launch_rigid_frame(Solver& s) {
  if (s.npieces <= 0) return;
  PieceArgs A = make_args(s);
  // per piece node: x, dx, state, mass, 12 rigidk planes (twice: centroid then E)
  ProfScope ps(s, PC_PIECES, s.N * 0.0);
  k_rigid_frame<<<piece_grid(s.npieces), kWG, 0, s.stream>>>(A);
}

void launch_rigid_start(Solver& s) {
  if (s.npieces <= 0) return;
  PieceArgs A = make_args(s);
  ProfScope ps(s, PC_PIECES, s.N * 0.0);
  k_rigid_reduce<<<piece_grid(s.npieces), kWG, 0, s.stream>>>(A);
  k_rigid_apply<<<grid_for(s.N, 256), 256, 0, s.stream>>>(A);
}

void launch_spin(Solver& s) {
  if (s.npieces <= 0) return;
  PieceArgs A = make_args(s);
  ProfScope ps(s, PC_PIECES, s.N * 0.0);
  k_spin_reduce<<<piece_grid(s.npieces), kWG, 0, s.stream>>>(A);
  k_spin_apply<<<grid_for(s.N, 256), 256, 0, s.stream>>>(A);
}

void launch_curve_set(Solver& s) {
  if (s.npieces <= 0) return;
  PieceArgs A = make_args(s);
  ProfScope ps(s, PC_PIECES, s.N * 0.0);
  k_curve_set<<<grid_for(s.N, 256), 256, 0, s.stream>>>(A);
}

void launch_curve_apply(Solver& s) {
  if (s.npieces <= 0) return;
  PieceArgs A = make_args(s);
  ProfScope ps(s, PC_PIECES, s.N * 0.0);
  k_curve_apply<<<grid_for(s.N, 256), 256, 0, s.stream>>>(A);
}

}  // namespace bd

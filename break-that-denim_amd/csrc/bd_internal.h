// Internal layout of a solver handle and the launchers each kernel file provides.
//
// HBM layout (structure of arrays; every plane padded to a multiple of 64 entries):
//   node vectors      3 planes of float  [x | y | z], e.g. x[c * Np + n]
//   fine operator     32 planes of float per node (128 B): the symmetric centre block (6), the
//                     forward symmetric blocks E (+1,0), N (0,+1), NE (+1,+1), NW (-1,+1) (6 each)
//                     and the along-line scalar couplings EE (+2,0), NN (0,+2). Every block of the
//                     clamp-projected Hessian is symmetric and A(m,n) = A(n,m), so backward
//                     couplings are read from the neighbour's forward planes.
//   coarse operators  13 symmetric blocks per node (78 planes, 312 B): centre plus the 12 forward
//                     offsets of the radius-2 stencil (Galerkin blocks stay symmetric).
//   prolongation      4 float weights per fine node, in bracket order (j outer, i inner).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bd.h"
#include "bd_math.h"

namespace bd {

constexpr int kFineComps = 32;
constexpr int kCoarseBlocks = 13;
constexpr int kCoarseComps = 6 * kCoarseBlocks;
constexpr int kLineColors = 3;  // line colours per family: lines of one colour share no coupling
// grid of the deterministic reductions: one 1024-thread block per CU
constexpr int kDotBlocks = 256;
constexpr int kDotThreads = 1024;
constexpr int kMaxLevels = BD_MAX_LEVELS;

// forward offsets of the coarse stencil, block b = 1..12 (block 0 is the centre)
__host__ __device__ constexpr int coarse_di(int b) {
  return b == 1 ? 1 : b == 2 ? 2 : b == 3 ? -2 : b == 4 ? -1 : b == 5 ? 0 : b == 6 ? 1 : b == 7 ? 2 : b == 8 ? -2 : b == 9 ? -1 : b == 10 ? 0 : b == 11 ? 1 : b == 12 ? 2 : 0;
}
__host__ __device__ constexpr int coarse_dj(int b) { return b == 0 ? 0 : b <= 2 ? 0 : b <= 7 ? 1 : 2; }
// block index of forward offset (di, dj), or -1 when (di, dj) is not a forward offset
__host__ __device__ inline int coarse_block(int di, int dj) {
  if (di < -2 || di > 2) return -1;
  if (dj == 0) return di == 0 ? 0 : di == 1 ? 1 : di == 2 ? 2 : -1;
  if (dj == 1) return 5 + di;   // di in [-2, 2] -> 3..7
  if (dj == 2) return 10 + di;  // -> 8..12
  return -1;
}

// fine forward-plane bases
constexpr int F_C = 0, F_E = 6, F_N = 12, F_NE = 18, F_NW = 24, F_EE = 30, F_NN = 31;

struct Level {
  int N = 0, Np = 0;
  Geo geo{};
  bool fine = false;
  float* op = nullptr;
  uint8_t* fixed = nullptr;
  float* x = nullptr;   // 3 planes
  float* b = nullptr;
  float* r = nullptr;
  float* pw = nullptr;  // 4 planes: prolongation weights to the next level
  // line tables: per colour (6: v-lines 0..2, u-lines 0..2) a list of int4 (start, stride, len, family)
  int4* lines = nullptr;
  int color_start[7] = {0};
  int ncol = 3;  // colours per family (v-lines 0 .. ncol-1, then u-lines): kLineColors
  int max_len = 0;
  // dual layout (lines of kDualMinLine nodes or more): column-major copies (node (i, j) of a layer at
  // off + i * ny + j) of the operator, fixed flags and iterate for the v-line passes; the v-line
  // entries of `lines` then start in that layout with stride 1
  bool dual = false;
  float* opT = nullptr;
  uint8_t* fixedT = nullptr;
  float* xT = nullptr;
  float* bT = nullptr;  // column-major copy of b, refreshed before the level is smoothed
  // dual levels with lines of kBatchCopyMinLine nodes or more: a pass writes x in its own layout only,
  // and one tiled copy after each family's passes refreshes the other layout (a line's result lies
  // strided in the other layout: written by every pass, it costs more than the pass's solve)
  bool batch_copy = false;
  // one-colour levels (bd_set_line_colors): the second buffer of the iterate in each layout; a sweep's
  // first pass writes them, its second pass writes x and xT back (allocated by the setter)
  float* x2 = nullptr;
  float* xT2 = nullptr;
  // smoother kernel per family (u, v): -1 the default rule, 0 direct, 4 / 8 staged, 16 split
  // (bd_autotune)
  int variant[2] = {-1, -1};
  // split variant (levels with lines longer than kStagedMaxLen): the assembled rows of one colour
  // pass, 18 planes (D 6, U 9, r 3) of rows_cap = lines per colour x longest line
  float* rows = nullptr;
  size_t rows_cap = 0;
  // line blocks of every line of the level (BD_SMOOTHER_SPLIT4 reads them): per row D (6) and U (9)
  // as the gather assembles them, which the iterate does not change, refreshed whenever the level's
  // operator is final; 15 planes of du_cap = lines of both families x the line stride
  float* du = nullptr;
  size_t du_cap = 0;
  int du_stride = 0;
  // the lines' factorizations (BD_SMOOTHER_FACTORED): per row its pivot inverse and elimination
  // multipliers (kFacPlanes = 24 planes of du_cap, same positions as du), refreshed with du
  float* fac = nullptr;
  // BD_PRECISION_FP16: op (and opT) as FP16 pairs, comps / 2 u32 planes (oprd.h); hs[0] the scale,
  // hs[1] its inverse, hs[2] the bits of the largest free diagonal entry
  uint32_t* oph = nullptr;
  uint32_t* ophT = nullptr;
  float* hs = nullptr;
};

// longest line the staged smoother kernel takes by default; longer lines use the direct kernel
constexpr int kStagedMaxLen = 257;

// levels whose longest line has at least this many nodes keep the column-major copy: below it a
// v-line pass's strided gather costs less than the copy (about twice a u-line pass at 65 nodes)
constexpr int kDualMinLine = 33;
// dual levels whose longest line has at least this many nodes refresh the other layout's copy of x
// once per family's passes instead of in every pass
constexpr int kBatchCopyMinLine = 512;
// longest line the one-workgroup PCG solves (its LDS holds the line records of a colour round)
constexpr int kWgMaxLine = 512;

struct Scal {
  // PCG and line-search scalars kept on the device
  double bb, rz, rz_new, pap, alpha, beta, rr, relres;
  double e[4], m[4];
  double line_alpha;
  double tol;
  int done, iters, max_iters, fixed_iters;
  int rejections, pad;
  int need_true, restart, last, replacements;
  // learned correction: on for this solve, switched off by the fallback, its gate for the
  // transfer kernels (done or off), fallbacks taken, the residual history of the fallback test
  int learn_on, learn_off, lgate, learn_fallbacks;
  double hist[8];
  double warm;               // relative residual of the start (1: zero start)
  double rec_warm[BD_MAX_NEWTON];
  unsigned int ticket;       // last-block-done counter of the current reduction
  unsigned int max_update[BD_MAX_NEWTON];  // float bits
  // per Newton iteration record of the substep
  int rec_iters[BD_MAX_NEWTON];
  double rec_relres[BD_MAX_NEWTON];
  double rec_alpha[BD_MAX_NEWTON];
};

// kernel classes for bd_profile_read
enum ProfClass {
  PC_PREDICT_COMMIT, PC_ASSEMBLE, PC_PROLONG_WEIGHTS, PC_RAP, PC_SMOOTH_FINE, PC_SMOOTH_COARSE, PC_RESIDUAL_FINE,
  PC_RESIDUAL_COARSE, PC_RESTRICT, PC_PROLONG, PC_MATVEC, PC_PCG_VECTOR, PC_TRUE_RESIDUAL, PC_ENERGY, PC_DAMAGE, PC_SLIP,
  PC_CONTACT_HASH, PC_CONTACT_DETECT, PC_CONTACT_TERMS, PC_TRANSPOSE, PC_COARSE_FACTOR, PC_COARSE_INVERSE, PC_COARSE_APPLY, PC_PIECES,
  PC_LEARNED, PC_COUNT
};
inline const char* prof_name(int c) {
  static const char* n[PC_COUNT] = {"predict_commit", "assemble", "prolong_weights", "rap", "smooth_fine", "smooth_coarse",
                                    "residual_fine", "residual_coarse", "restrict", "prolong", "matvec_dot", "pcg_vector",
                                    "true_residual", "energy", "damage", "slip", "contact_hash", "contact_detect",
                                    "contact_terms", "transpose", "coarse_factor", "coarse_inverse", "coarse_apply", "pieces",
                                    "learned"};
  return n[c];
}

struct ProfRec {
  int cls;
  int ev;  // index of the start event in the pool (stop = ev + 1)
  double bytes;
};

// a grip slot (interact.wgsl GripSlot): handle target in the frame of `anchor`, entry count,
// stiffness per unit weight, largest lead, picked node, pinch scale, pad pitch, slip distance
struct GripSlot {
  float goal[3];
  int count;
  float stiffness, max_lead;
  uint32_t picked;
  float pinch;
  float anchor[3];
  float pad_pitch;
  float slip;
  float pad_[3];
};
// a patch entry (interact.wgsl GripEntry): node, weight, pad side (-1, +1, 0 radial), grab-time offset
struct GripEntry {
  uint32_t node;
  float w, side, pad0;
  float offset[3];
  float pad1;
};
// a pick query (interact.wgsl PickU)
struct PickArgs {
  float view_proj[16];  // column-major
  float pointer[2];
  float radius_px, width, height;
  float strand_radius;  // strand pick radius (m); the search radius of a nearest-material pick
  int window_radius;
  float eye[3], ray[3];
};
constexpr int kGrips = BD_MAX_GRIPS_HELD;
constexpr int kSlotEntries = BD_GRIP_SLOT_ENTRIES;

struct Solver {
  bd_params P{};
  // profiling
  bool prof_on = false;
  std::vector<hipEvent_t> prof_pool;
  std::vector<ProfRec> prof_recs;
  int N = 0, Np = 0, C = 0;
  int nl = 0;
  DevLayer layers_h[8];
  DevLayer* layers = nullptr;  // device copy
  Geo geo{};
  hipStream_t stream = nullptr;
  // side stream of the exact coarsest factorization (default priority, like `stream`): it starts when
  // the hierarchy is built and the V-cycle waits for it only at its first coarse solve, so the
  // factorization (one workgroup, then the inverse's 55) runs beside the fine and coarse smoothing
  hipStream_t aux = nullptr;
  hipEvent_t ev_rap = nullptr, ev_fact = nullptr;
  bool fact_pending = false;
  std::string err;

  // node state (3 planes each)
  float *x = nullptr, *v = nullptr, *dx = nullptr, *pred = nullptr;
  double* delta = nullptr;  // Newton direction accumulated in FP64 (iterative refinement)
  float *rhs = nullptr, *p = nullptr, *ap = nullptr;   // PCG; r and z are level 0's b and x
  float* last_update = nullptr;
  float* mass = nullptr;
  double* cn0 = nullptr;
  uint8_t* pinned = nullptr;
  // node state (state.wgsl): 0 cloth, 1 detached (a piece no mount or grip reaches), 2 retired (left
  // the scene: frozen, spans removed). `pinned` holds bit kPinMount for mounted nodes and bit
  // kPinRetired for retired ones, so every kernel that skips pinned rows skips retired ones too.
  std::vector<uint8_t> mount_h;  // host copy of the mounted flags (topology anchors)
  // grips and picking (interact.wgsl)
  GripSlot* grip_slots = nullptr;
  GripSlot grip_slots_h[BD_MAX_GRIPS_HELD] = {};
  float grip_full_k[BD_MAX_GRIPS_HELD] = {};
  GripEntry* grip_entries = nullptr;
  float* grip_handle = nullptr;  // 12 per grip: effective handle, origin, closing axis
  bool grips_active = false;
  uint32_t* pick = nullptr;
  uint32_t* pick_out = nullptr;
  PickArgs pick_args{};
  uint8_t* nstate = nullptr;
  bool any_detached = false;  // some node has been detached (commit checks for retirement)
  // detached pieces and their rigid modes (rigid_frame/apply/spin.wgsl)
  int npieces = 0;
  uint32_t* piece_start = nullptr;  // kMaxPieces + 1: first entry of piece p in piece_nodes
  uint32_t* piece_nodes = nullptr;  // N
  uint32_t* piece_of = nullptr;     // Np: piece of each node, or kNoPiece
  double* piece_frame = nullptr;    // kPieceFrame per piece: centroid xyz, valid, E^-1 (6 x 6)
  double* piece_step = nullptr;     // 8 per piece: rigid step t (or spin centroid), rotation w
  float* rigidk = nullptr;          // 12 planes: external block, internal rotation stiffness
  float* curve = nullptr;           // 7 planes: rotation w xyz, flag, arm r xyz
  uint8_t* layer_of = nullptr;
  // spans
  double *rest = nullptr, *intact = nullptr, *fail_site = nullptr, *tails = nullptr, *softening = nullptr;
  uint32_t* cut_mask = nullptr;
  uint16_t* flags = nullptr;
  double* strain = nullptr;
  // damage field
  uint8_t *damage = nullptr, *area = nullptr, *sep = nullptr, *pullmark = nullptr;
  float* strength = nullptr;
  // grips
  int grip_count = 0;
  int* grip_of = nullptr;      // per node
  float* grip_w = nullptr;     // per node
  double* grip_off = nullptr;  // 3 planes
  double* grip_goal_k = nullptr;  // per grip: goal xyz, k
  // pinned staging ring for per-substep goal updates (a slot is reused after its copy completed)
  static constexpr int kGoalSlots = 16;
  double* grip_goal_h = nullptr;     // kGoalSlots x 4 * BD_MAX_GRIPS
  hipEvent_t goal_ev[kGoalSlots] = {};
  bool goal_ev_used[kGoalSlots] = {};
  int goal_slot = 0;
  // reductions
  double* partials = nullptr;  // kDotBlocks * 8
  Scal* scal = nullptr;
  Scal* scal_h = nullptr;      // pinned host mirror
  int* flag_h = nullptr;       // coherent pinned: PCG's convergence flag [0] and its sequence number [1]
  int* flag_d = nullptr;       // the same memory as the device addresses it
  int flag_seq = 0;
  int poll_timeout_ms = 2000;  // BD_POLL_TIMEOUT_MS at creation; 0 takes the synchronizing path at once
  int poll_fallbacks = 0;      // convergence checks that took the synchronizing path
  // learned coarse-space correction (learned.hip): network, recorded samples, chain vectors
  struct Learn* learn = nullptr;
  int learned = 0;             // BD_LEARNED_*: the learned warm start, and the deflated correction
  // graph-captured substep (bd_graph_capture): while capturing, graph_pcg caps each Newton solve's PCG
  // iterations and replaces the host's convergence poll by the device gate (iterations queued past
  // convergence return at once); the graph, its executable and its Newton count; config_gen counts
  // the changes of what a substep launches, graph_gen is its value at the capture
  int graph_pcg = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_newton = 0;
  unsigned config_gen = 0, graph_gen = 0;
  unsigned graph_serial = 0;  // captures so far (a batch keeps the serial of the graph it copied)
  // PCG path (bd_set_pcg_mode): BD_PCG_GRID kernels over the chip, or BD_PCG_WORKGROUP, the whole
  // solve in one workgroup (wgsolve.hip) with its device descriptor and dynamic LDS
  int pcg_mode = BD_PCG_GRID;
  void* wg_sys = nullptr;
  size_t wg_lds = 0;
  int* counters = nullptr;     // [0] ruptures, [1] pulled out
  // multigrid
  int nlev = 0;
  Level lv[kMaxLevels];
  bool assembled = false, hierarchy = false;
  bd_stats stats{};
  int coarse_valid_iteration = -1;

  // contact: self-contact (node against intact cell) of a one-layer scene, cross-layer contact
  // (node against node of another ply) of a scene of several layers, and the floor
  int contact_mode = 0;  // kContactNone, kContactSelf or kContactPlies
  bool floor_on = false;
  float floor_z = 0.f;
  uint32_t* ijl = nullptr;      // per node: i | j << 12 | layer << 24
  uint32_t* cells = nullptr;    // 2 * kTable: counts scanned into starts, then fill cursors
  uint32_t* scan_sums = nullptr;
  float4* items = nullptr;      // 2 per record, at most 2 records per face
  uint32_t* pairs = nullptr;    // kContactK planes of Np: candidate faces
  unsigned long long* keys = nullptr;  // reciprocal-bound keys, share code << 32 | node + 1 (cleared by resolve)
  uint32_t* marks = nullptr;    // damaged-row marks
  float4* cmeta = nullptr;      // 2 per node: own bound, reciprocal bound
  int* acc = nullptr;           // 9 planes of fixed-point accumulators
  float* ext = nullptr;         // 9 planes: gradient xyz, Hessian xx xy xz yy yz zz
  bool contact_detected = false;
  // cross-layer contact: a (cell, layer) node hash of 2 x 2^20 u32 (counts scanned into starts,
  // then fill cursors), node ids in cell order, the K nearest foreign nodes in `pairs`
  uint32_t* pcells = nullptr;
  uint32_t* psums = nullptr;     // 1024 block sums of the scan + 1
  uint32_t* pitems = nullptr;    // N node ids
  float ply_dist[3] = {0.f, 0.f, 0.f};  // contact distance of layer pairs 0-1, 0-2, 1-2 (m)

  // Newton iteration of the substep under way (0 outside substeps): BD_COARSE_EXACT_LAGGED factors
  // the coarsest level only at iteration 0
  int newton_it = 0;
  bool coarse_factored = false;
  // exact coarsest solve (BD_COARSE_EXACT, BD_COARSE_EXACT_LAGGED): banded Cholesky factor and dense
  // inverse, FP64
  int cn = 0, cbw = 0;
  double* cband = nullptr;
  double* cinv = nullptr;
  double* crd = nullptr;   // reciprocal pivots of the factor
};

constexpr int kContactNone = 0, kContactSelf = 1, kContactPlies = 2;
constexpr uint8_t kStateCloth = 0, kStateDetached = 1, kStateRetired = 2;
constexpr uint8_t kPinMount = 1, kPinRetired = 2;
constexpr int kMaxPieces = 4096;
constexpr uint32_t kNoPiece = 0xFFFFFFFFu;
constexpr int kPieceFrame = 40;
constexpr int kContactK = 4;
constexpr int kPlyTableBits = 20;
constexpr int kPlyTable = 1 << kPlyTableBits;
constexpr int kMaxPlies = 3;
constexpr int kTableBits = 17;
constexpr int kTable = 1 << kTableBits;

// launchers (assemble.hip)
void launch_predict(Solver& s);
void launch_assemble(Solver& s);
void launch_newton_apply(Solver& s, int it);
void launch_commit(Solver& s);
// multigrid.hip
void launch_prolong_weights(Solver& s, int k);
void launch_rap(Solver& s, int k);
void launch_residual(Solver& s, int k, const float* x, const float* b, float* r);
void launch_restrict(Solver& s, int k);
void launch_prolong_add(Solver& s, int k);
void launch_smooth(Solver& s, int k, int reverse);
std::vector<int> smooth_variants(Solver& s, int k, int fam);
void smooth_alloc_rows(Solver& s, Level& L);
void smooth_line_blocks(Solver& s, int k);
void launch_zero(Solver& s, float* p, size_t n);
// column-major copies of a dual-layout level: operator and fixed flags, or the iterate
void launch_colmajor_op(Solver& s, int k);
void launch_colmajor_x(Solver& s, int k);
// x of level k in one layout copied to the other: row-major -> column-major (to_col) or back
void launch_layout_copy(Solver& s, int k, const float* src, float* dst, bool to_col);
void launch_colmajor_b(Solver& s, int k);
// a vector's restriction to level k + 1 (xc of the transfer untouched) and prolongation from it,
// both gated by `gate` (skipped while it is nonzero); `set`: the fine vector is overwritten, else added to
void launch_restrict_vec(Solver& s, int k, const float* r, float* bc, const int* gate);
// y = A_k x and x^T A_k x (FP64) into device memory; scratch: kApplyScratch bytes of device memory
constexpr size_t kApplyScratch = (kDotBlocks + 2) * sizeof(double);
void launch_apply_dot(Solver& s, int k, const float* x, float* y, double* scratch, double* xax, hipStream_t st);
// y = A_k x with FP64 products and sums, rounded once to FP32
void launch_apply_f64(Solver& s, int k, const float* x, float* y);
void launch_prolong_vec(Solver& s, int k, const float* xc, float* xf, bool set, const int* gate);
void vcycle(Solver& s, int k);
// pcg.hip
void pcg_solve(Solver& s);
void launch_pcg_setup(Solver& s);
std::string wg_unsupported(const Solver& s);
void wg_prepare(Solver& s);
void wg_free(Solver& s);
void launch_pcg_wg(Solver& s);
size_t wg_sys_bytes();
void launch_pcg_wg_many(const void* systems, int count, size_t lds, hipStream_t st);
inline const int* pcg_done(Solver& s) { return &s.scal->done; }
inline const int* learn_off(Solver& s) { return &s.scal->learn_off; }
inline const int* learn_gate(Solver& s) { return &s.scal->lgate; }
// learned.hip
void learn_alloc(Solver& s);
void learn_free(Solver& s);
int learn_level(const Solver& s);
void learn_predict(Solver& s);
void learn_correct(Solver& s, const float* r, float* z);
void learn_deflate(Solver& s, float* z);
bool learn_ready(const Solver& s);
void learn_record_start(Solver& s, int capacity);
bool learn_recording(const Solver& s);
void learn_record_begin(Solver& s, const float* r);
void learn_record_mid(Solver& s, const float* r, const double* delta);
void learn_record_end(Solver& s, const double* delta, const float* rhs);
int learn_sample_count(const Solver& s);
void learn_train(Solver& s, int epochs, double lr, int batch, int targets, double* losses);
void learn_loss_grad(Solver& s, int sample, double* loss, float* grad);
void learn_get_weights(Solver& s, float* out);
void learn_set_weights(Solver& s, const float* w);
void learn_read_sample(Solver& s, int sample, float* b, double* st, float* op, uint8_t* fixed);
void learn_read_z(Solver& s, float* z);
int learn_padded(const Solver& s);
int learn_nodes(const Solver& s);
void launch_energy(Solver& s);
// history.hip
void launch_damage(Solver& s);
void launch_slip(Solver& s);
// contact.hip
void contact_alloc(Solver& s);
void contact_reset_bounds(Solver& s);
void launch_contact_detect(Solver& s);
void launch_contact_terms(Solver& s);
void launch_project_start(Solver& s);
// plies.hip
void ply_contact_alloc(Solver& s);
void launch_ply_detect(Solver& s);
void launch_ply_terms(Solver& s);
// pieces.hip
void pieces_alloc(Solver& s);
void launch_apply_state(Solver& s);
void launch_rigid_frame(Solver& s);
void launch_rigid_start(Solver& s);
void launch_spin(Solver& s);
void launch_curve_set(Solver& s);
void launch_curve_apply(Solver& s);
// interact.hip
void interact_alloc(Solver& s);
void launch_grip_init(Solver& s, int g);
void launch_grip_lead(Solver& s);
void launch_grip_terms(Solver& s);
void launch_pick(Solver& s, bool ray);
// coarse.hip
void coarse_exact_alloc(Solver& s);
// the V-cycle solves its coarsest level exactly (BD_COARSE_EXACT or BD_COARSE_EXACT_LAGGED)
inline bool coarse_exact(const Solver& s) {
  return s.P.coarse_solve == BD_COARSE_EXACT || s.P.coarse_solve == BD_COARSE_EXACT_LAGGED;
}
void launch_coarse_factor(Solver& s);
void launch_coarse_apply(Solver& s);

#define BD_HIP(call)                                                     \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

inline int grid_for(int n, int block) { return (n + block - 1) / block; }

// Brackets the launches in its scope with HIP events while profiling is on.
// the solver's stream waits for a coarsest factorization still running on the side stream
inline void coarse_join(Solver& s) {
  if (!s.fact_pending) return;
  BD_HIP(hipStreamWaitEvent(s.stream, s.ev_fact, 0));
  s.fact_pending = false;
}

struct ProfScope {
  Solver& s;
  int rec = -1;
  ProfScope(Solver& s_, int cls, double bytes) : s(s_) {
    if (!s.prof_on) return;
    int ev = (int)s.prof_recs.size() * 2;
    while ((int)s.prof_pool.size() < ev + 2) {
      hipEvent_t e;
      BD_HIP(hipEventCreate(&e));
      s.prof_pool.push_back(e);
    }
    BD_HIP(hipEventRecord(s.prof_pool[ev], s.stream));
    s.prof_recs.push_back(ProfRec{cls, ev, bytes});
    rec = (int)s.prof_recs.size() - 1;
  }
  ~ProfScope() {
    if (rec >= 0) (void)hipEventRecord(s.prof_pool[s.prof_recs[rec].ev + 1], s.stream);
  }
};

}  // namespace bd

// C-ABI of libbd.so (include/bd.h): solver handles, uploads, the substep schedule, stage entry
// points and readbacks. Stands in for the reference's GpuSolver (src/simulation/gpu/solver.ts).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <thread>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "bd_internal.h"

struct bd_solver {
  bd::Solver s;
};

namespace bd {
void smooth_set_attributes();
void launch_record(Solver& s, int it);
void launch_open_gate(Solver& s);
void launch_reset_counters(Solver& s);
}  // namespace bd

using namespace bd;

static thread_local std::string g_create_error;

// Scratch of bd_apply_noalloc: the FP64 block partials of its reduction (kDotBlocks doubles), its
// ticket and its result slot. It does not grow with the level (the reduction grid is capped at
// kDotBlocks blocks). Every device the library runs on holds its own copy until the process ends,
// and every call on that device shares it: calls that overlap on one device race.
static __device__ __align__(256) double scratch[kDotBlocks + 2];
static_assert(sizeof(scratch) >= kApplyScratch, "bd_apply_noalloc scratch");
static std::mutex g_scratch_mu;
static std::vector<double*> g_scratch_ptr;  // per device: the array's device address

// the scratch array's address on the current device, looked up once per device (the array's host
// name is only a stand-in for it)
static hipError_t resolve_scratch(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

namespace {

template <class T>
T* dalloc(size_t count) {
  if (count == 0) count = 1;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(T));
  if (e != hipSuccess) throw std::runtime_error(std::string("hipMalloc: ") + hipGetErrorString(e));
  // hipMemset runs on the null stream and may return before it completes; the solver's stream is
  // non-blocking and is not ordered after it, so an upload or kernel issued there next could land
  // first and then be cleared. Wait for the clear here.
  BD_HIP(hipMemset(p, 0, count * sizeof(T)));
  BD_HIP(hipDeviceSynchronize());
  return static_cast<T*>(p);
}

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

int pad64(int n) { return (n + 63) / 64 * 64; }

// interleaved xyz doubles -> 3 float planes
void upload_vec3(Solver& s, float* dst, int N, int Np, const double* src) {
  std::vector<float> h((size_t)3 * Np, 0.f);
  for (int n = 0; n < N; n++)
    for (int q = 0; q < 3; q++) h[(size_t)q * Np + n] = (float)src[3 * n + q];
  BD_HIP(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, s.stream));
  BD_HIP(hipStreamSynchronize(s.stream));
}

void download_vec3(Solver& s, const float* src, int N, int Np, double* dst) {
  std::vector<float> h((size_t)3 * Np);
  BD_HIP(hipMemcpyAsync(h.data(), src, h.size() * sizeof(float), hipMemcpyDeviceToHost, s.stream));
  BD_HIP(hipStreamSynchronize(s.stream));
  for (int n = 0; n < N; n++)
    for (int q = 0; q < 3; q++) dst[3 * n + q] = h[(size_t)q * Np + n];
}

template <class T>
void upload(Solver& s, T* dst, const T* src, size_t count) {
  BD_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, s.stream));
  BD_HIP(hipStreamSynchronize(s.stream));
}

template <class T>
void download(Solver& s, T* dst, const T* src, size_t count) {
  BD_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, s.stream));
  BD_HIP(hipStreamSynchronize(s.stream));
}

DevLayer dev_layer(const bd_layer& L) {
  DevLayer d{};
  d.nx = L.nx;
  d.ny = L.ny;
  d.off = L.node_offset;
  d.bu = L.bundle_u;
  d.bv = L.bundle_v;
  d.foff = L.field_offset;
  d.fk = L.field_k;
  d.fm = L.field_m;
  d.wpu = L.weave_pitch_u;
  d.wpv = L.weave_pitch_v;
  d.pu = L.material.pitch_u;
  d.pv = L.material.pitch_v;
  for (int f = 0; f < 2; f++) {
    const bd_yarn& y = L.material.yarns[f];
    d.y[f] = DevYarn{y.linear_density, y.ea_low, y.ea_high, y.crimp_strain, y.ea_compress, y.axial_viscosity, y.bend_rigidity,
                     y.bend_viscosity, y.strength, y.softening_opening, y.yarns_per_span, 0};
  }
  d.shear_k = L.material.shear_stiffness;
  d.mu = L.material.crossing_friction;
  d.cn0 = L.material.crossing_normal0;
  d.wrap = L.material.crossing_wrap;
  d.air_drag = L.material.air_drag;
  d.thickness = L.material.thickness;
  return d;
}

// level layouts (lattice.ts:43-66 coarsenLayers / levelLayers, newtonMultigrid.ts:480-491)
std::vector<Geo> level_layouts(const Geo& fine, int max_coarsest) {
  std::vector<Geo> out{fine};
  for (;;) {
    const Geo& cur = out.back();
    int N = 0;
    bool can = true;
    for (int l = 0; l < cur.nl; l++) {
      N += cur.nx[l] * cur.ny[l];
      if (!(cur.nx[l] * cur.ny[l] > 48 && cur.nx[l] >= 5 && cur.ny[l] >= 5)) can = false;
    }
    if (N <= max_coarsest || !can || (int)out.size() >= kMaxLevels) break;
    Geo c{};
    c.nl = cur.nl;
    int off = 0;
    for (int l = 0; l < cur.nl; l++) {
      c.nx[l] = cur.nx[l] / 2 + 1;  // ceil((n - 1) / 2) + 1
      c.ny[l] = cur.ny[l] / 2 + 1;
      c.off[l] = off;
      off += c.nx[l] * c.ny[l];
    }
    out.push_back(c);
  }
  return out;
}

// line tables (lattice.ts:86-106 lineTable): per colour the lines of one family, v-lines (family 1)
// first; `ncol` colours per family (kLineColors = 3 by default: the fine level's along-line bending
// couples lines two apart as well, so lines of one colour three apart share no coupling; with 1 colour
// (bd_set_line_colors) every line of a family is solved at once against the lagged iterate of the
// others)
void build_lines(Solver& s, Level& L, int ncol) {
  L.ncol = ncol;
  std::vector<std::vector<int4>> by(2 * ncol);
  int maxlen = 0;
  for (int l = 0; l < L.geo.nl; l++) {
    int nx = L.geo.nx[l], ny = L.geo.ny[l], off = L.geo.off[l];
    for (int i = 0; i < nx; i++)
      by[i % ncol].push_back(L.dual ? make_int4(off + i * ny, 1, ny, 1) : make_int4(off + i, nx, ny, 1));
    for (int j = 0; j < ny; j++) by[ncol + j % ncol].push_back(make_int4(off + j * nx, 1, nx, 0));
    maxlen = std::max(maxlen, std::max(nx, ny));
  }
  std::vector<int4> flat;
  L.color_start[0] = 0;
  for (int c = 0; c < 2 * ncol; c++) {
    flat.insert(flat.end(), by[c].begin(), by[c].end());
    L.color_start[c + 1] = (int)flat.size();
  }
  L.max_len = maxlen;
  L.lines = dalloc<int4>(flat.size());
  upload(s, L.lines, flat.data(), flat.size());
}

// the captured substep, if any
void release_graph(Solver& s) {
  if (s.graph_exec) (void)hipGraphExecDestroy(s.graph_exec);
  if (s.graph) (void)hipGraphDestroy(s.graph);
  s.graph_exec = nullptr;
  s.graph = nullptr;
}

void destroy_all(Solver& s) {
  // nothing may still run on either stream while the buffers go
  if (s.aux) (void)hipStreamSynchronize(s.aux);
  if (s.stream) (void)hipStreamSynchronize(s.stream);
  void* ptrs[] = {s.layers, s.x, s.v, s.dx, s.pred, s.delta, s.rhs, s.p, s.ap, s.last_update, s.mass, s.cn0, s.pinned,
                  s.layer_of, s.rest, s.intact, s.fail_site, s.tails, s.softening, s.cut_mask, s.flags, s.strain,
                  s.damage, s.area, s.sep, s.pullmark, s.strength, s.grip_of, s.grip_w, s.grip_off, s.grip_goal_k,
                  s.partials, s.scal, s.counters, s.ijl, s.cells, s.scan_sums, s.items, s.pairs, s.keys,
                  s.marks, s.cmeta, s.acc, s.ext, s.cband, s.cinv, s.crd, s.pcells, s.psums, s.pitems,
                  s.nstate, s.piece_start, s.piece_nodes, s.piece_of, s.piece_frame, s.piece_step, s.rigidk, s.curve,
                  s.grip_slots, s.grip_entries, s.grip_handle, s.pick, s.pick_out};
  for (void* p : ptrs) dfree(p);
  for (int k = 0; k < s.nlev; k++) {
    Level& L = s.lv[k];
    dfree(L.op);
    if (k > 0) dfree(L.fixed);  // level 0 shares the solver's pinned flags
    dfree(L.x);
    dfree(L.b);
    dfree(L.r);
    dfree(L.pw);
    dfree(L.lines);
    dfree(L.opT);
    dfree(L.fixedT);
    dfree(L.xT);
    dfree(L.bT);
    dfree(L.rows);
    dfree(L.du);
    dfree(L.fac);
    dfree(L.x2);
    dfree(L.xT2);
    dfree(L.oph);
    dfree(L.ophT);
    dfree(L.hs);
  }
  if (s.scal_h) (void)hipHostFree(s.scal_h);
  if (s.flag_h) (void)hipHostFree(s.flag_h);
  if (s.grip_goal_h) (void)hipHostFree(s.grip_goal_h);
  for (int q = 0; q < Solver::kGoalSlots; q++)
    if (s.goal_ev[q]) (void)hipEventDestroy(s.goal_ev[q]);
  for (hipEvent_t e : s.prof_pool) (void)hipEventDestroy(e);
  learn_free(s);
  s.prof_pool.clear();
  release_graph(s);
  wg_free(s);
  if (s.aux) (void)hipStreamSynchronize(s.aux);
  if (s.aux) (void)hipStreamDestroy(s.aux);
  if (s.ev_rap) (void)hipEventDestroy(s.ev_rap);
  if (s.ev_fact) (void)hipEventDestroy(s.ev_fact);
  if (s.stream) (void)hipStreamDestroy(s.stream);
}

int fail(bd_solver* h, int code, const std::string& msg) {
  if (h) h->s.err = msg;
  return code;
}

template <class F>
int guarded(bd_solver* h, F&& f) {
  if (!h) return BD_ERR_INVALID;
  try {
    return f(h->s);
  } catch (const std::exception& e) {
    return fail(h, BD_ERR_HIP, e.what());
  }
}

void build_hierarchy(Solver& s) {
  for (int k = 0; k + 1 < s.nlev; k++) {
    launch_prolong_weights(s, k);
    launch_rap(s, k);
    launch_colmajor_op(s, k + 1);
  }
  // the lagged mode keeps the factorization of the substep's first Newton iteration
  if (coarse_exact(s) && (s.P.coarse_solve == BD_COARSE_EXACT || s.newton_it == 0 || !s.coarse_factored)) {
    // on the side stream, after everything queued so far (the RAP it reads, the V-cycles that read the
    // previous factorization); its kernels and profile events go there while the streams are swapped
    coarse_join(s);
    BD_HIP(hipEventRecord(s.ev_rap, s.stream));
    BD_HIP(hipStreamWaitEvent(s.aux, s.ev_rap, 0));
    std::swap(s.stream, s.aux);
    launch_coarse_factor(s);
    std::swap(s.stream, s.aux);
    BD_HIP(hipEventRecord(s.ev_fact, s.aux));
    s.fact_pending = true;
    s.coarse_factored = true;
  }
  s.hierarchy = true;
  if (s.learned) learn_predict(s);
}

void contact_terms(Solver& s) {
  if (s.contact_mode == kContactSelf) launch_contact_terms(s);
  else if (s.contact_mode == kContactPlies) launch_ply_terms(s);
}

void contact_detect(Solver& s) {
  if (s.contact_mode == kContactSelf) launch_contact_detect(s);
  else if (s.contact_mode == kContactPlies) launch_ply_detect(s);
}

// encodeAssemble (gpu/solver.ts:613-622): the Newton system, the detached pieces' rigid modes at
// this iterate, and the Galerkin hierarchy when asked for
void assemble_system(Solver& s, bool coarse) {
  launch_assemble(s);
  launch_rigid_frame(s);
  s.assembled = true;
  if (coarse || !s.hierarchy) build_hierarchy(s);
}

// encodePredict (gpu/solver.ts:697-705): the inertial prediction, and detached pieces start from
// their rotation over the substep
void predict(Solver& s) {
  launch_predict(s);
  launch_spin(s);
}

// the grip slot on the device
void upload_slot(Solver& s, int g) {
  BD_HIP(hipMemcpyAsync(s.grip_slots + g, &s.grip_slots_h[g], sizeof(GripSlot), hipMemcpyHostToDevice, s.stream));
  BD_HIP(hipStreamSynchronize(s.stream));
  bool any = false;
  for (int q = 0; q < kGrips; q++) any = any || s.grip_slots_h[q].count > 0;
  if (any != s.grips_active) s.config_gen++;
  s.grips_active = any;
}

void fill_pick(Solver& s, int window_radius, bd_pick* out) {
  std::vector<uint32_t> w(8 + 256);
  download(s, w.data(), s.pick_out, w.size());
  float p[3];
  std::memcpy(p, w.data(), sizeof(p));
  out->kind = (int32_t)w[4];
  out->node = w[4] == 0 ? -1 : (int32_t)w[3];
  for (int q = 0; q < 3; q++) out->position[q] = p[q];
  out->window_radius = window_radius;
  const int side = 2 * window_radius + 1;
  std::memset(out->window, 0, sizeof(out->window));
  for (int k = 0; k < side * side; k++) out->window[k] = w[8 + k];
}

// encodeLineSearch (gpu/solver.ts:678-684)
void line_search(Solver& s) {
  launch_curve_set(s);
  launch_energy(s);
}

// a Newton iteration before its linear solve: contact and grip terms at this iterate (grip terms
// added to contact's, GpuSolver.preAssemble, interaction.ts:341-345), the system and its hierarchy
void newton_pre(Solver& s, bool coarse) {
  contact_terms(s);
  launch_grip_terms(s);
  assemble_system(s, coarse);
}

// and after it: line search, the update, the iteration's record
void newton_post(Solver& s, int it) {
  line_search(s);
  launch_curve_apply(s);
  launch_newton_apply(s, it);
  launch_record(s, it);
}

void newton_iteration(Solver& s, int it, bool coarse) {
  s.newton_it = it;
  newton_pre(s, coarse);
  pcg_solve(s);
  newton_post(s, it);
}

// one substep (encodeSubstep, gpu/solver.ts:737-760): prediction, grip lead, contact detection and
// projectStart, n_it Newton iterations and the optional ones up to n_max while the last update moved a
// node by more than the Newton tolerance (gate.wgsl), commit, damage field and slip; returns the Newton
// iterations run
void substep_begin(Solver& s) {
  launch_reset_counters(s);
  predict(s);
  // once per substep: each grip's handle lead over its patch (interaction.ts:335-338)
  launch_grip_lead(s);
  contact_detect(s);
  if (s.contact_mode != kContactNone || s.floor_on) launch_project_start(s);
}

void substep_end(Solver& s) {
  s.newton_it = 0;
  coarse_join(s);
  launch_commit(s);
  if (s.P.history) {
    launch_damage(s);
    launch_slip(s);
  }
}

int substep_body(Solver& s, int n_it, int n_max) {
  substep_begin(s);
  int done = 0;
  for (int it = 0; it < n_max; it++) {
    if (it >= n_it) {
      // gated optional iteration (gate.wgsl): only while the last update moved a node > tolerance
      BD_HIP(hipMemcpyAsync(s.scal_h->max_update, s.scal->max_update, sizeof(s.scal_h->max_update), hipMemcpyDeviceToHost, s.stream));
      BD_HIP(hipStreamSynchronize(s.stream));
      float moved;
      std::memcpy(&moved, &s.scal_h->max_update[it - 1], sizeof(float));
      if (!(moved > (float)s.P.newton_tolerance)) break;
    }
    newton_iteration(s, it, it == 0 || !s.P.lag_coarse);
    done = it + 1;
  }
  substep_end(s);
  return done;
}

void ensure_pieces(Solver& s) {
  if (!s.piece_start) pieces_alloc(s);
}

}  // namespace

extern "C" {

const char* bd_last_error(const bd_solver* h) { return h ? h->s.err.c_str() : g_create_error.c_str(); }

int bd_create(const bd_scene* sc, const bd_params* prm, bd_solver** out) {
  g_create_error.clear();
  if (!sc || !prm || !out) {
    g_create_error = "bd_create: null argument";
    return BD_ERR_INVALID;
  }
  *out = nullptr;
  if (sc->node_count <= 0 || sc->layer_count <= 0 || sc->layer_count > 8) {
    g_create_error = "bd_create: need 1..8 layers and a positive node count";
    return BD_ERR_INVALID;
  }
  int total = 0;
  for (int l = 0; l < sc->layer_count; l++) {
    const bd_layer& L = sc->layers[l];
    if (L.nx < 2 || L.ny < 2 || L.node_offset != total) {
      g_create_error = "bd_create: layers must be at least 2x2 and stored contiguously";
      return BD_ERR_INVALID;
    }
    if (L.nx > BD_MAX_LINE || L.ny > BD_MAX_LINE) {
      g_create_error = "bd_create: yarn line longer than BD_MAX_LINE (" + std::to_string(BD_MAX_LINE) + ")";
      return BD_ERR_LIMIT;
    }
    if (L.nx > 4096 || L.ny > 4096) {
      g_create_error = "bd_create: lattice index beyond 12 bits";
      return BD_ERR_LIMIT;
    }
    if (L.bundle_u < 1 || L.bundle_v < 1 || L.bundle_u > 32 || L.bundle_v > 32) {
      g_create_error = "bd_create: bundle too large for the cut masks";
      return BD_ERR_LIMIT;
    }
    total += L.nx * L.ny;
  }
  if (total != sc->node_count) {
    g_create_error = "bd_create: node_count does not match the layers";
    return BD_ERR_INVALID;
  }
  if (prm->coarse_solve != BD_COARSE_SWEEPS && prm->coarse_solve != BD_COARSE_EXACT &&
      prm->coarse_solve != BD_COARSE_EXACT_LAGGED) {
    g_create_error = "bd_create: unknown coarse_solve";
    return BD_ERR_INVALID;
  }
  if (prm->precision != BD_PRECISION_FP32 && prm->precision != BD_PRECISION_FP16) {
    g_create_error = "bd_create: unknown precision";
    return BD_ERR_INVALID;
  }
  if (prm->newton_iterations < 1 || prm->newton_iterations > BD_MAX_NEWTON) {
    g_create_error = "bd_create: newton_iterations must be 1.." + std::to_string(BD_MAX_NEWTON);
    return BD_ERR_INVALID;
  }
  if (prm->smooth_sweeps < 1 || prm->smooth_sweeps > 8) {
    g_create_error = "bd_create: smooth_sweeps must be 1..8";
    return BD_ERR_INVALID;
  }
  if (sc->layer_count > BD_MAX_CONTACT_LAYERS) {
    g_create_error = "bd_create: cross-layer contact supports at most 3 layers";
    return BD_ERR_LIMIT;
  }
  if (prm->self_contact && sc->layer_count == 1 && sc->node_count > BD_CONTACT_MAX_NODES) {
    g_create_error = "bd_create: self-contact supports at most BD_CONTACT_MAX_NODES nodes";
    return BD_ERR_LIMIT;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    g_create_error = "bd_create: no HIP device";
    return BD_ERR_HIP;
  }
  {
    // set-up of bd_apply_noalloc's scratch on this device
    double* sp = nullptr;
    const hipError_t e = resolve_scratch(&sp);
    if (e != hipSuccess) {
      g_create_error = std::string("bd_create: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return BD_ERR_HIP;
    }
  }
  bd_solver* h = new bd_solver();
  Solver& s = h->s;
  try {
    s.P = *prm;
    s.N = sc->node_count;
    s.Np = pad64(s.N);
    s.C = sc->crossing_count;
    s.nl = sc->layer_count;
    BD_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    BD_HIP(hipStreamCreateWithFlags(&s.aux, hipStreamNonBlocking));
    BD_HIP(hipEventCreateWithFlags(&s.ev_rap, hipEventDisableTiming));
    BD_HIP(hipEventCreateWithFlags(&s.ev_fact, hipEventDisableTiming));
    s.geo.nl = s.nl;
    for (int l = 0; l < s.nl; l++) {
      s.layers_h[l] = dev_layer(sc->layers[l]);
      s.geo.nx[l] = sc->layers[l].nx;
      s.geo.ny[l] = sc->layers[l].ny;
      s.geo.off[l] = sc->layers[l].node_offset;
    }
    s.layers = dalloc<DevLayer>(s.nl);
    upload(s, s.layers, s.layers_h, s.nl);
    const int N = s.N, Np = s.Np;
    s.x = dalloc<float>(3 * (size_t)Np);
    s.v = dalloc<float>(3 * (size_t)Np);
    s.dx = dalloc<float>(3 * (size_t)Np);
    s.pred = dalloc<float>(3 * (size_t)Np);
    s.delta = dalloc<double>(3 * (size_t)Np);
    s.rhs = dalloc<float>(3 * (size_t)Np);
    s.p = dalloc<float>(3 * (size_t)Np);
    s.ap = dalloc<float>(3 * (size_t)Np);
    s.last_update = dalloc<float>(Np);
    s.mass = dalloc<float>(Np);
    s.cn0 = dalloc<double>(Np);
    s.pinned = dalloc<uint8_t>(Np);
    s.rest = dalloc<double>(2 * (size_t)N);
    s.intact = dalloc<double>(2 * (size_t)N);
    s.fail_site = dalloc<double>(2 * (size_t)N);
    s.tails = dalloc<double>(4 * (size_t)N);
    s.softening = dalloc<double>(2 * (size_t)N);
    s.strain = dalloc<double>(2 * (size_t)N);
    s.cut_mask = dalloc<uint32_t>(2 * (size_t)N);
    s.flags = dalloc<uint16_t>(2 * (size_t)N);
    s.damage = dalloc<uint8_t>(2 * (size_t)s.C);
    s.area = dalloc<uint8_t>(2 * (size_t)s.C);
    s.sep = dalloc<uint8_t>(2 * (size_t)s.C);
    s.pullmark = dalloc<uint8_t>(2 * (size_t)s.C);
    s.strength = dalloc<float>(2 * (size_t)s.C);
    s.grip_of = dalloc<int>(Np);
    s.grip_w = dalloc<float>(Np);
    s.grip_off = dalloc<double>(3 * (size_t)Np);
    s.grip_goal_k = dalloc<double>(4 * BD_MAX_GRIPS);
    s.partials = dalloc<double>(kDotBlocks * 8);
    s.scal = dalloc<Scal>(1);
    s.counters = dalloc<int>(8);
    BD_HIP(hipHostMalloc((void**)&s.scal_h, sizeof(Scal), hipHostMallocDefault));
    std::memset(s.scal_h, 0, sizeof(Scal));
    BD_HIP(hipHostMalloc((void**)&s.flag_h, 2 * sizeof(int), hipHostMallocCoherent | hipHostMallocMapped));
    s.flag_h[0] = s.flag_h[1] = 0;
    BD_HIP(hipHostGetDevicePointer((void**)&s.flag_d, s.flag_h, 0));
    if (const char* t = std::getenv("BD_POLL_TIMEOUT_MS")) s.poll_timeout_ms = std::max(0, std::atoi(t));
    const size_t goal_bytes = sizeof(double) * 4 * BD_MAX_GRIPS * (Solver::kGoalSlots + 1);
    BD_HIP(hipHostMalloc((void**)&s.grip_goal_h, goal_bytes, hipHostMallocDefault));
    std::memset(s.grip_goal_h, 0, goal_bytes);
    for (int q = 0; q < Solver::kGoalSlots; q++) BD_HIP(hipEventCreateWithFlags(&s.goal_ev[q], hipEventDisableTiming));

    std::vector<float> f(Np, 0.f);
    for (int n = 0; n < N; n++) f[n] = (float)sc->mass[n];
    upload(s, s.mass, f.data(), Np);
    upload(s, s.cn0, sc->crossing_normal0, N);
    {
      // bit kPinMount: mounted by the scene; bit kPinRetired is set when a node retires
      s.mount_h.assign(N, 0);
      for (int n = 0; n < N; n++) s.mount_h[n] = sc->pinned[n] ? kPinMount : 0;
      upload(s, s.pinned, s.mount_h.data(), N);
      s.nstate = dalloc<uint8_t>(Np);
      BD_HIP(hipMemsetAsync(s.nstate, 0, Np, s.stream));
    }
    std::vector<int> gof(Np, -1);
    upload(s, s.grip_of, gof.data(), Np);
    upload(s, s.rest, sc->rest_length, 2 * (size_t)N);
    upload(s, s.intact, sc->intact, 2 * (size_t)N);
    upload(s, s.fail_site, sc->fail_site, 2 * (size_t)N);
    upload(s, s.tails, sc->tails, 4 * (size_t)N);
    upload(s, s.softening, sc->softening, 2 * (size_t)N);
    upload(s, s.cut_mask, sc->cut_mask, 2 * (size_t)N);
    upload(s, s.flags, sc->seg_flags, 2 * (size_t)N);
    if (s.C > 0) {
      upload(s, s.damage, sc->field_damage, 2 * (size_t)s.C);
      upload(s, s.area, sc->field_area, 2 * (size_t)s.C);
      upload(s, s.strength, sc->field_strength, 2 * (size_t)s.C);
    }
    upload_vec3(s, s.x, N, Np, sc->position);

    // contact: lattice coordinates, external terms (zero unless self-contact writes them), bounds
    {
      std::vector<uint32_t> ijl(Np, 0u);
      for (int l = 0; l < s.nl; l++) {
        const bd_layer& L = sc->layers[l];
        for (int j = 0; j < L.ny; j++)
          for (int i = 0; i < L.nx; i++) ijl[L.node_offset + j * L.nx + i] = (uint32_t)i | ((uint32_t)j << 12) | ((uint32_t)l << 24);
      }
      s.ijl = dalloc<uint32_t>(Np);
      upload(s, s.ijl, ijl.data(), Np);
      s.ext = dalloc<float>(9 * (size_t)Np);
      BD_HIP(hipMemsetAsync(s.ext, 0, 9 * (size_t)Np * sizeof(float), s.stream));
      s.cmeta = dalloc<float4>(2 * (size_t)Np);
      contact_reset_bounds(s);
      // (the grip and pick buffers are allocated on the first attach or pick)
    // a scene of several plies always has cross-layer contact; a one-layer scene self-contact
      // when asked for (gpu/solver.ts:322, 364)
      s.contact_mode = s.nl > 1 ? kContactPlies : (s.P.self_contact ? kContactSelf : kContactNone);
      if (s.contact_mode == kContactSelf) contact_alloc(s);
      if (s.contact_mode == kContactPlies) ply_contact_alloc(s);
      BD_HIP(hipStreamSynchronize(s.stream));
    }

    std::vector<Geo> geos = level_layouts(s.geo, prm->max_coarsest_nodes);
    s.nlev = (int)geos.size();
    for (int k = 0; k < s.nlev; k++) {
      Level& L = s.lv[k];
      L.geo = geos[k];
      L.N = 0;
      for (int l = 0; l < L.geo.nl; l++) L.N += L.geo.nx[l] * L.geo.ny[l];
      L.Np = pad64(L.N);
      L.fine = k == 0;
      L.op = dalloc<float>((size_t)(L.fine ? kFineComps : kCoarseComps) * L.Np);
      if (k == 0) {
        L.fixed = s.pinned;
        L.x = dalloc<float>(3 * (size_t)L.Np);  // PCG z
        L.b = dalloc<float>(3 * (size_t)L.Np);  // PCG r
      } else {
        L.fixed = dalloc<uint8_t>(L.Np);
        L.x = dalloc<float>(3 * (size_t)L.Np);
        L.b = dalloc<float>(3 * (size_t)L.Np);
      }
      L.r = dalloc<float>(3 * (size_t)L.Np);
      if (k + 1 < s.nlev) L.pw = dalloc<float>(4 * (size_t)L.Np);
      int longest = 0;
      for (int l = 0; l < L.geo.nl; l++) longest = std::max(longest, std::max(L.geo.nx[l], L.geo.ny[l]));
      L.dual = longest >= kDualMinLine;
      L.batch_copy = L.dual && longest >= kBatchCopyMinLine;
      if (L.dual) {
        L.opT = dalloc<float>((size_t)(L.fine ? kFineComps : kCoarseComps) * L.Np);
        L.fixedT = dalloc<uint8_t>(L.Np);
        L.xT = dalloc<float>(3 * (size_t)L.Np);
        BD_HIP(hipMemsetAsync(L.xT, 0, 3 * (size_t)L.Np * sizeof(float), s.stream));
        L.bT = dalloc<float>(3 * (size_t)L.Np);
      }
      build_lines(s, L, kLineColors);
      if (L.max_len > kStagedMaxLen) smooth_alloc_rows(s, L);
      // FP16 on the Galerkin levels only: FP16 fine operators cost 15% more PCG iterations at 1024^2, the
      // coarse ones none (and save 4% of the frame)
      if (s.P.precision == BD_PRECISION_FP16 && !L.fine) {
        const size_t words = (size_t)(L.fine ? kFineComps : kCoarseComps) / 2 * L.Np;
        L.oph = dalloc<uint32_t>(words);
        if (L.dual) L.ophT = dalloc<uint32_t>(words);
        L.hs = dalloc<float>(4);
      }
    }
    smooth_set_attributes();
    if (coarse_exact(s)) coarse_exact_alloc(s);
    s.stats.level_count = s.nlev;
    for (int k = 0; k < s.nlev && k < BD_MAX_LEVELS; k++) s.stats.level_nodes[k] = s.lv[k].N;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    destroy_all(s);
    delete h;
    return BD_ERR_HIP;
  }
  *out = h;
  return BD_OK;
}

int bd_destroy(bd_solver* h) {
  if (!h) return BD_OK;
  (void)hipStreamSynchronize(h->s.stream);
  destroy_all(h->s);
  delete h;
  return BD_OK;
}

int bd_set_state(bd_solver* h, const double* x, const double* v) {
  return guarded(h, [&](Solver& s) {
    if (!x) return fail(h, BD_ERR_INVALID, "bd_set_state: null positions");
    upload_vec3(s, s.x, s.N, s.Np, x);
    if (v)
      upload_vec3(s, s.v, s.N, s.Np, v);
    else
      BD_HIP(hipMemsetAsync(s.v, 0, 3 * (size_t)s.Np * sizeof(float), s.stream));
    BD_HIP(hipStreamSynchronize(s.stream));
    return BD_OK;
  });
}

int bd_set_segments(bd_solver* h, const double* rest, const double* intact, const uint32_t* cut, const uint16_t* flags,
                    const double* fail_site, const double* tails) {
  return guarded(h, [&](Solver& s) {
    size_t n2 = 2 * (size_t)s.N;
    if (rest) upload(s, s.rest, rest, n2);
    if (intact) upload(s, s.intact, intact, n2);
    if (cut) upload(s, s.cut_mask, cut, n2);
    if (flags) upload(s, s.flags, flags, n2);
    if (fail_site) upload(s, s.fail_site, fail_site, n2);
    if (tails) upload(s, s.tails, tails, 2 * n2);
    return BD_OK;
  });
}

int bd_set_field(bd_solver* h, const uint8_t* damage) {
  return guarded(h, [&](Solver& s) {
    if (!damage) return fail(h, BD_ERR_INVALID, "bd_set_field: null damage");
    if (s.C > 0) upload(s, s.damage, damage, 2 * (size_t)s.C);
    return BD_OK;
  });
}

int bd_set_grips(bd_solver* h, const bd_grips* g) {
  return guarded(h, [&](Solver& s) {
    std::vector<int> gof(s.Np, -1);
    std::vector<float> w(s.Np, 0.f);
    std::vector<double> off(3 * (size_t)s.Np, 0.0), gk(4 * BD_MAX_GRIPS, 0.0);
    int count = 0;
    if (g) {
      if (g->grip_count < 0 || g->grip_count > BD_MAX_GRIPS) return fail(h, BD_ERR_LIMIT, "bd_set_grips: more than BD_MAX_GRIPS grips");
      count = g->grip_count;
      for (int k = 0; k < count; k++) {
        for (int q = 0; q < 3; q++) gk[4 * k + q] = g->goal[3 * k + q];
        gk[4 * k + 3] = g->stiffness[k];
      }
      for (int e = 0; e < g->entry_count; e++) {
        int n = g->node[e];
        if (n < 0 || n >= s.N || g->grip[e] < 0 || g->grip[e] >= count)
          return fail(h, BD_ERR_INVALID, "bd_set_grips: entry out of range");
        gof[n] = g->grip[e];
        w[n] = (float)g->weight[e];
        for (int q = 0; q < 3; q++) off[(size_t)q * s.Np + n] = g->offset[3 * e + q];
      }
    }
    s.grip_count = count;
    s.config_gen++;
    std::memcpy(s.grip_goal_h + (size_t)Solver::kGoalSlots * 4 * BD_MAX_GRIPS, gk.data(), sizeof(double) * 4 * BD_MAX_GRIPS);
    upload(s, s.grip_of, gof.data(), gof.size());
    upload(s, s.grip_w, w.data(), w.size());
    upload(s, s.grip_off, off.data(), off.size());
    upload(s, s.grip_goal_k, gk.data(), gk.size());
    return BD_OK;
  });
}

int bd_set_grip_goals(bd_solver* h, int count, const double* goals) {
  return guarded(h, [&](Solver& s) {
    if (count != s.grip_count || (count > 0 && !goals)) return fail(h, BD_ERR_INVALID, "bd_set_grip_goals: grip count differs from bd_set_grips");
    const int slot = s.goal_slot;
    s.goal_slot = (slot + 1) % Solver::kGoalSlots;
    if (s.goal_ev_used[slot]) BD_HIP(hipEventSynchronize(s.goal_ev[slot]));
    double* hbuf = s.grip_goal_h + (size_t)slot * 4 * BD_MAX_GRIPS;
    std::memcpy(hbuf, s.grip_goal_h + (size_t)Solver::kGoalSlots * 4 * BD_MAX_GRIPS, sizeof(double) * 4 * BD_MAX_GRIPS);
    for (int k = 0; k < count; k++)
      for (int q = 0; q < 3; q++) hbuf[4 * k + q] = goals[3 * k + q];
    std::memcpy(s.grip_goal_h + (size_t)Solver::kGoalSlots * 4 * BD_MAX_GRIPS, hbuf, sizeof(double) * 4 * BD_MAX_GRIPS);
    BD_HIP(hipMemcpyAsync(s.grip_goal_k, hbuf, sizeof(double) * 4 * BD_MAX_GRIPS, hipMemcpyHostToDevice, s.stream));
    BD_HIP(hipEventRecord(s.goal_ev[slot], s.stream));
    s.goal_ev_used[slot] = true;
    return BD_OK;
  });
}

int bd_set_timestep(bd_solver* h, double dt) {
  return guarded(h, [&](Solver& s) {
    if (!(dt > 0)) return fail(h, BD_ERR_INVALID, "bd_set_timestep: h must be positive");
    s.P.h = dt;
    s.config_gen++;
    return BD_OK;
  });
}

int bd_set_history(bd_solver* h, int on) {
  return guarded(h, [&](Solver& s) {
    s.P.history = on ? 1 : 0;
    s.config_gen++;
    return BD_OK;
  });
}

int bd_set_newton_tolerance(bd_solver* h, double tol) {
  return guarded(h, [&](Solver& s) {
    s.P.newton_tolerance = tol;
    s.config_gen++;
    return BD_OK;
  });
}

int bd_set_pcg(bd_solver* h, int iterations, double rel_tol, int max_iterations) {
  return guarded(h, [&](Solver& s) {
    if (iterations < 0 || rel_tol < 0 || (rel_tol > 0 && max_iterations < 1)) return fail(h, BD_ERR_INVALID, "bd_set_pcg: bad schedule");
    s.P.pcg_iterations = iterations;
    s.P.pcg_rel_tol = rel_tol;
    s.P.pcg_max_iterations = max_iterations;
    s.config_gen++;
    return BD_OK;
  });
}

int bd_set_pcg_mode(bd_solver* h, int mode) {
  return guarded(h, [&](Solver& s) {
    if (mode != BD_PCG_GRID && mode != BD_PCG_WORKGROUP) return fail(h, BD_ERR_INVALID, "bd_set_pcg_mode: unknown mode");
    if (mode == BD_PCG_WORKGROUP) {
      const std::string why = wg_unsupported(s);
      if (!why.empty()) return fail(h, BD_ERR_STATE, "bd_set_pcg_mode: no one-workgroup solve: " + why);
      BD_HIP(hipStreamSynchronize(s.stream));
      wg_prepare(s);
    }
    s.pcg_mode = mode;
    s.config_gen++;
    return BD_OK;
  });
}

int bd_predict(bd_solver* h) {
  return guarded(h, [&](Solver& s) {
    predict(s);
    return BD_OK;
  });
}

int bd_assemble(bd_solver* h, int coarse) {
  return guarded(h, [&](Solver& s) {
    assemble_system(s, coarse != 0);
    return BD_OK;
  });
}

int bd_vcycle(bd_solver* h) {
  return guarded(h, [&](Solver& s) {
    if (!s.hierarchy) return fail(h, BD_ERR_STATE, "bd_vcycle: assemble first");
    launch_open_gate(s);
    launch_zero(s, s.lv[0].x, 3 * (size_t)s.lv[0].Np);
    if (s.lv[0].dual) launch_zero(s, s.lv[0].xT, 3 * (size_t)s.lv[0].Np);
    vcycle(s, 0);
    return BD_OK;
  });
}

int bd_smooth(bd_solver* h, int level, int reverse) {
  return guarded(h, [&](Solver& s) {
    if (!s.hierarchy) return fail(h, BD_ERR_STATE, "bd_smooth: assemble first");
    if (level < 0 || level >= s.nlev) return fail(h, BD_ERR_INVALID, "bd_smooth: level out of range");
    launch_open_gate(s);
    launch_colmajor_x(s, level);
    launch_colmajor_b(s, level);
    launch_smooth(s, level, reverse);
    return BD_OK;
  });
}

int bd_residual(bd_solver* h, int level) {
  return guarded(h, [&](Solver& s) {
    if (!s.hierarchy) return fail(h, BD_ERR_STATE, "bd_residual: assemble first");
    if (level < 0 || level >= s.nlev) return fail(h, BD_ERR_INVALID, "bd_residual: level out of range");
    launch_open_gate(s);
    Level& L = s.lv[level];
    launch_residual(s, level, L.x, L.b, L.r);
    return BD_OK;
  });
}

int bd_pcg(bd_solver* h) {
  return guarded(h, [&](Solver& s) {
    if (!s.hierarchy) return fail(h, BD_ERR_STATE, "bd_pcg: assemble first");
    pcg_solve(s);
    launch_record(s, 0);
    return BD_OK;
  });
}

int bd_line_search(bd_solver* h) {
  return guarded(h, [&](Solver& s) {
    line_search(s);
    return BD_OK;
  });
}

int bd_newton_apply(bd_solver* h, int it) {
  return guarded(h, [&](Solver& s) {
    if (it < 0 || it >= BD_MAX_NEWTON) return fail(h, BD_ERR_INVALID, "bd_newton_apply: iteration out of range");
    launch_curve_apply(s);
    launch_newton_apply(s, it);
    return BD_OK;
  });
}

int bd_finalize(bd_solver* h) {
  return guarded(h, [&](Solver& s) {
    launch_commit(s);
    if (s.P.history) {
      launch_damage(s);
      launch_slip(s);
    }
    return BD_OK;
  });
}

int bd_damage(bd_solver* h) {
  return guarded(h, [&](Solver& s) {
    launch_damage(s);
    return BD_OK;
  });
}

int bd_slip(bd_solver* h) {
  return guarded(h, [&](Solver& s) {
    launch_slip(s);
    return BD_OK;
  });
}

int bd_contact_detect(bd_solver* h) {
  return guarded(h, [&](Solver& s) {
    if (s.contact_mode == kContactNone)
      return fail(h, BD_ERR_STATE, "bd_contact_detect: one-layer scene created without self_contact");
    contact_detect(s);
    return BD_OK;
  });
}

int bd_contact_terms(bd_solver* h) {
  return guarded(h, [&](Solver& s) {
    if (s.contact_mode == kContactNone)
      return fail(h, BD_ERR_STATE, "bd_contact_terms: one-layer scene created without self_contact");
    contact_terms(s);
    return BD_OK;
  });
}

int bd_project_start(bd_solver* h) {
  return guarded(h, [&](Solver& s) {
    launch_project_start(s);
    return BD_OK;
  });
}

int bd_set_floor(bd_solver* h, int enabled, double z) {
  return guarded(h, [&](Solver& s) {
    BD_HIP(hipStreamSynchronize(s.stream));
    s.floor_on = enabled != 0;
    s.floor_z = enabled ? (float)z : 0.f;
    s.config_gen++;
    return BD_OK;
  });
}

int bd_grip_attach(bd_solver* h, int slot, int32_t picked, const int32_t* nodes, const float* weights, const float* sides,
                   int count, double stiffness, double max_lead, double pad_pitch, double slip, const double* anchor) {
  return guarded(h, [&](Solver& s) {
    if (slot < 0 || slot >= kGrips) return fail(h, BD_ERR_INVALID, "bd_grip_attach: slot out of range");
    if (picked < 0 || picked >= s.N || count < 0 || (count > 0 && (!nodes || !weights || !sides)) || !anchor)
      return fail(h, BD_ERR_INVALID, "bd_grip_attach: bad arguments");
    // entries beyond a slot's capacity are dropped (interaction.ts:228)
    const int c = count < kSlotEntries ? count : kSlotEntries;
    std::vector<GripEntry> e(c);
    for (int k = 0; k < c; k++) {
      if (nodes[k] < 0 || nodes[k] >= s.N) return fail(h, BD_ERR_INVALID, "bd_grip_attach: node out of range");
      e[k] = GripEntry{(uint32_t)nodes[k], weights[k], sides[k], 0.f, {0.f, 0.f, 0.f}, 0.f};
    }
    interact_alloc(s);
    if (c > 0) upload(s, s.grip_entries + (size_t)slot * kSlotEntries, e.data(), (size_t)c);
    GripSlot& S = s.grip_slots_h[slot];
    S = GripSlot{};
    for (int q = 0; q < 3; q++) {
      S.goal[q] = (float)anchor[q];
      S.anchor[q] = (float)anchor[q];
    }
    S.count = c;
    S.stiffness = (float)stiffness;
    S.max_lead = (float)max_lead;
    S.picked = (uint32_t)picked;
    S.pinch = 1.f;
    S.pad_pitch = (float)pad_pitch;
    S.slip = (float)slip;
    s.grip_full_k[slot] = (float)stiffness;
    upload_slot(s, slot);
    launch_grip_init(s, slot);
    return BD_OK;
  });
}

int bd_grip_lead(bd_solver* h) {
  return guarded(h, [&](Solver& s) {
    launch_grip_lead(s);
    return BD_OK;
  });
}

int bd_grip_terms(bd_solver* h) {
  return guarded(h, [&](Solver& s) {
    launch_grip_terms(s);
    return BD_OK;
  });
}

int bd_grip_target(bd_solver* h, int slot, const double* target) {
  return guarded(h, [&](Solver& s) {
    if (slot < 0 || slot >= kGrips || !target) return fail(h, BD_ERR_INVALID, "bd_grip_target: bad arguments");
    if (!s.grip_slots || s.grip_slots_h[slot].count == 0) return BD_OK;
    for (int q = 0; q < 3; q++) s.grip_slots_h[slot].goal[q] = (float)target[q];
    upload_slot(s, slot);
    return BD_OK;
  });
}

int bd_grip_pinch(bd_solver* h, int slot, double scale) {
  return guarded(h, [&](Solver& s) {
    if (slot < 0 || slot >= kGrips) return fail(h, BD_ERR_INVALID, "bd_grip_pinch: slot out of range");
    if (!s.grip_slots || s.grip_slots_h[slot].count == 0) return BD_OK;
    s.grip_slots_h[slot].pinch = (float)scale;
    upload_slot(s, slot);
    return BD_OK;
  });
}

int bd_grip_hold(bd_solver* h, int slot, double fraction) {
  return guarded(h, [&](Solver& s) {
    if (slot < 0 || slot >= kGrips) return fail(h, BD_ERR_INVALID, "bd_grip_hold: slot out of range");
    if (!s.grip_slots || s.grip_slots_h[slot].count == 0) return BD_OK;
    s.grip_slots_h[slot].stiffness = (float)(s.grip_full_k[slot] * fraction);
    upload_slot(s, slot);
    return BD_OK;
  });
}

int bd_grip_release(bd_solver* h, int slot) {
  return guarded(h, [&](Solver& s) {
    if (slot < 0 || slot >= kGrips) return fail(h, BD_ERR_INVALID, "bd_grip_release: slot out of range");
    if (!s.grip_slots || s.grip_slots_h[slot].count == 0) return BD_OK;
    s.grip_slots_h[slot].count = 0;
    upload_slot(s, slot);
    // without contact the grips own ext: clear it (the remaining grips rewrite their entries)
    if (s.contact_mode == kContactNone) BD_HIP(hipMemsetAsync(s.ext, 0, 9 * (size_t)s.Np * sizeof(float), s.stream));
    return BD_OK;
  });
}

int bd_pick_ray(bd_solver* h, const float* view_proj, double px, double py, double width, double height,
                double radius_px, double strand_radius, int window_radius, const double* eye, const double* dir,
                bd_pick* out) {
  return guarded(h, [&](Solver& s) {
    if (!view_proj || !eye || !dir || !out || window_radius < 0 || window_radius > BD_PICK_MAX_WINDOW_RADIUS)
      return fail(h, BD_ERR_INVALID, "bd_pick_ray: bad arguments");
    interact_alloc(s);
    PickArgs& P = s.pick_args;
    for (int q = 0; q < 16; q++) P.view_proj[q] = view_proj[q];
    P.pointer[0] = (float)px;
    P.pointer[1] = (float)py;
    P.radius_px = (float)radius_px;
    P.width = (float)width;
    P.height = (float)height;
    P.strand_radius = (float)strand_radius;
    P.window_radius = window_radius;
    const double dl = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    if (!(dl > 0)) return fail(h, BD_ERR_INVALID, "bd_pick_ray: zero ray direction");
    for (int q = 0; q < 3; q++) {
      P.eye[q] = (float)eye[q];
      P.ray[q] = (float)(dir[q] / dl);  // unit: hit distances in metres
    }
    launch_pick(s, true);
    fill_pick(s, window_radius, out);
    return BD_OK;
  });
}

int bd_pick_near(bd_solver* h, const double* p, double radius, int window_radius, bd_pick* out) {
  return guarded(h, [&](Solver& s) {
    if (!p || !out || !(radius > 0) || window_radius < 0 || window_radius > BD_PICK_MAX_WINDOW_RADIUS)
      return fail(h, BD_ERR_INVALID, "bd_pick_near: bad arguments");
    interact_alloc(s);
    PickArgs& P = s.pick_args;
    P = PickArgs{};
    P.width = P.height = 1.f;
    P.strand_radius = (float)radius;
    P.window_radius = window_radius;
    for (int q = 0; q < 3; q++) P.eye[q] = (float)p[q];
    launch_pick(s, false);
    fill_pick(s, window_radius, out);
    return BD_OK;
  });
}

int bd_set_line_colors(bd_solver* h, int level, int colors) {
  return guarded(h, [&](Solver& s) {
    if (level < 0 || level >= s.nlev) return fail(h, BD_ERR_INVALID, "bd_set_line_colors: no such level");
    if (colors < 1 || colors > kLineColors) return fail(h, BD_ERR_INVALID, "bd_set_line_colors: colours must be 1, 2 or 3");
    if (s.pcg_mode == BD_PCG_WORKGROUP && colors != kLineColors)
      return fail(h, BD_ERR_STATE, "bd_set_line_colors: the one-workgroup PCG smooths with 3 colours");
    Level& L = s.lv[level];
    if (colors == 2 && !L.rows)
      return fail(h, BD_ERR_INVALID, "bd_set_line_colors: 2 colours need the split kernels (lines of more than 257 nodes)");
    if (L.ncol == colors) return BD_OK;
    coarse_join(s);
    BD_HIP(hipStreamSynchronize(s.stream));
    // two colours run only the split kernels
    if (colors == 2)
      for (int& v : L.variant)
        if (v != BD_SMOOTHER_SPLIT && v != BD_SMOOTHER_SPLIT4 && v != BD_SMOOTHER_FACTORED) v = BD_SMOOTHER_FACTORED;
    dfree(L.lines);
    L.lines = nullptr;
    build_lines(s, L, colors);
    if (colors == 1 && !L.x2) {
      L.x2 = dalloc<float>(3 * (size_t)L.Np);
      if (L.dual) L.xT2 = dalloc<float>(3 * (size_t)L.Np);
    }
    if (L.rows) {
      // the split variants' buffers are laid out by colour: rebuilt, and the line blocks refilled if
      // the level's operator is already there
      smooth_alloc_rows(s, L);
      if (level == 0 ? s.assembled : s.hierarchy) smooth_line_blocks(s, level);
    }
    s.config_gen++;
    return BD_OK;
  });
}

int bd_set_smoother(bd_solver* h, int level, int family, int variant) {
  return guarded(h, [&](Solver& s) {
    if (!s.hierarchy) return fail(h, BD_ERR_STATE, "bd_set_smoother: assemble first");
    if (level < 0 || level >= s.nlev) return fail(h, BD_ERR_INVALID, "bd_set_smoother: no such level");
    if (family != 0 && family != 1) return fail(h, BD_ERR_INVALID, "bd_set_smoother: family must be 0 (u) or 1 (v)");
    if (variant != BD_SMOOTHER_DEFAULT) {
      std::vector<int> vs = smooth_variants(s, level, family);
      if (std::find(vs.begin(), vs.end(), variant) == vs.end())
        return fail(h, BD_ERR_INVALID, "bd_set_smoother: the level cannot run that variant");
    }
    s.lv[level].variant[family] = variant;
    s.config_gen++;
    return BD_OK;
  });
}

int bd_autotune(bd_solver* h, int reps, int32_t* chosen) {
  return guarded(h, [&](Solver& s) {
    if (!s.hierarchy) return fail(h, BD_ERR_STATE, "bd_autotune: assemble first");
    if (reps < 1) return fail(h, BD_ERR_INVALID, "bd_autotune: reps must be positive");
    // Greedy over levels and families, fine first: each variant is timed inside whole V-cycles
    // (the others at their current choice), so that the level's data is as warm or cold as in the
    // solve; the fastest of `reps` V-cycles counts. b = x = 0: every kernel still does all its work.
    hipEvent_t e0, e1;
    BD_HIP(hipEventCreate(&e0));
    BD_HIP(hipEventCreate(&e1));
    Level& F = s.lv[0];
    auto time_vcycle = [&]() {
      launch_zero(s, F.b, 3 * (size_t)F.Np);
      launch_zero(s, F.x, 3 * (size_t)F.Np);
      if (F.dual) launch_zero(s, F.xT, 3 * (size_t)F.Np);
      BD_HIP(hipEventRecord(e0, s.stream));
      vcycle(s, 0);
      BD_HIP(hipEventRecord(e1, s.stream));
      BD_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      BD_HIP(hipEventElapsedTime(&ms, e0, e1));
      return ms;
    };
    launch_open_gate(s);
    for (int k = 0; k < s.nlev; k++)
      for (int fam = 1; fam >= 0; fam--) {
        std::vector<int> vs = smooth_variants(s, k, fam);
        if (vs.size() < 2) {
          s.lv[k].variant[fam] = vs.empty() ? -1 : vs[0];
          continue;
        }
        // the candidates in turn, `reps` rounds after a warm-up round, so that a slow stretch of the
        // device (clocks, other work) does not fall on one candidate's runs only; the fastest run of
        // each candidate counts
        std::vector<float> best(vs.size(), 0.f);
        for (int r = 0; r < reps + 1; r++)
          for (size_t c = 0; c < vs.size(); c++) {
            s.lv[k].variant[fam] = vs[c];
            const float ms = time_vcycle();
            if (r > 0 && (r == 1 || ms < best[c])) best[c] = ms;
          }
        size_t pick = 0;
        for (size_t c = 1; c < vs.size(); c++)
          if (best[c] < best[pick]) pick = c;
        s.lv[k].variant[fam] = vs[pick];
      }
    BD_HIP(hipEventDestroy(e0));
    BD_HIP(hipEventDestroy(e1));
    s.config_gen++;
    if (chosen)
      for (int k = 0; k < s.nlev; k++) {
        chosen[2 * k] = s.lv[k].variant[0];
        chosen[2 * k + 1] = s.lv[k].variant[1];
      }
    return BD_OK;
  });
}

int bd_set_pieces(bd_solver* h, const uint32_t* piece_of, const uint32_t* start, int count, const uint32_t* nodes) {
  return guarded(h, [&](Solver& s) {
    if (count < 0 || (count > 0 && (!piece_of || !start || !nodes)))
      return fail(h, BD_ERR_INVALID, "bd_set_pieces: bad arguments");
    // pieces beyond BD_MAX_PIECES go without the rigid-mode correction (gpu/solver.ts:664-676)
    const int c = count < kMaxPieces ? count : kMaxPieces;
    if (c > 0) {
      if (start[0] != 0) return fail(h, BD_ERR_INVALID, "bd_set_pieces: start[0] must be 0");
      for (int p = 0; p < c; p++)
        if (start[p + 1] < start[p]) return fail(h, BD_ERR_INVALID, "bd_set_pieces: start must not decrease");
      if (start[c] > (uint32_t)s.N) return fail(h, BD_ERR_INVALID, "bd_set_pieces: more piece entries than nodes");
      for (uint32_t k = 0; k < start[c]; k++)
        if (nodes[k] >= (uint32_t)s.N) return fail(h, BD_ERR_INVALID, "bd_set_pieces: node index out of range");
    }
    ensure_pieces(s);
    BD_HIP(hipStreamSynchronize(s.stream));
    std::vector<uint32_t> of(s.Np, kNoPiece);
    if (c > 0)
      for (int n = 0; n < s.N; n++) of[n] = piece_of[n] < (uint32_t)c ? piece_of[n] : kNoPiece;
    upload(s, s.piece_of, of.data(), (size_t)s.Np);
    if (c > 0) {
      upload(s, s.piece_start, start, (size_t)c + 1);
      if (start[c] > 0) upload(s, s.piece_nodes, nodes, (size_t)start[c]);
    }
    // no piece left: no node steps along a rotation
    if (c == 0 && s.npieces > 0) BD_HIP(hipMemsetAsync(s.curve, 0, 7 * (size_t)s.Np * sizeof(float), s.stream));
    s.npieces = c;
    s.config_gen++;
    BD_HIP(hipStreamSynchronize(s.stream));
    return BD_OK;
  });
}

int bd_update_topology(bd_solver* h, const int32_t* anchors, int anchor_count, int32_t* freed_out) {
  return guarded(h, [&](Solver& s) {
    if (anchor_count < 0 || (anchor_count > 0 && !anchors)) return fail(h, BD_ERR_INVALID, "bd_update_topology: bad anchors");
    const int N = s.N;
    std::vector<uint16_t> flags(2 * (size_t)N);
    std::vector<uint8_t> st(N);
    download(s, flags.data(), s.flags, 2 * (size_t)N);
    download(s, st.data(), s.nstate, (size_t)N);
    // pack (state.wgsl:16-27)
    std::vector<uint8_t> bits(N), freed(N);
    auto act = [&](size_t q) { return (flags[q] & (SEG_EXISTS | SEG_BROKEN)) == SEG_EXISTS; };
    for (int n = 0; n < N; n++)
      bits[n] = (uint8_t)((act(2 * (size_t)n) ? 1 : 0) | (act(2 * (size_t)n + 1) ? 2 : 0) | (st[n] != kStateCloth ? 4 : 0) |
                          (st[n] == kStateRetired ? 8 : 0));
    std::vector<int32_t> nx(s.nl), ny(s.nl), off(s.nl);
    for (int l = 0; l < s.nl; l++) {
      nx[l] = s.geo.nx[l];
      ny[l] = s.geo.ny[l];
      off[l] = s.geo.off[l];
    }
    std::vector<uint32_t> of(N), start((size_t)N + 1), nodes((size_t)N + 1);
    int32_t count = 0;
    const int nfree = bd_topology_search(N, s.nl, nx.data(), ny.data(), off.data(), bits.data(), s.mount_h.data(), anchors,
                                         anchor_count, freed.data(), of.data(), start.data(), nodes.data(), &count);
    if (nfree < 0) return fail(h, BD_ERR_INVALID, "bd_update_topology: anchors out of range");
    if (nfree > 0) {
      // applyFreed (state.wgsl:29-35): freed nodes become detached, irreversibly
      for (int n = 0; n < N; n++)
        if (freed[n] && st[n] == kStateCloth) st[n] = kStateDetached;
      upload(s, s.nstate, st.data(), (size_t)N);
      s.any_detached = true;
      s.config_gen++;
    }
    if (freed_out) *freed_out = nfree;
    return bd_set_pieces(h, of.data(), start.data(), count, nodes.data());
  });
}

int bd_substep(bd_solver* h, int min_it, int max_it) {
  return guarded(h, [&](Solver& s) {
    const int n_it = std::min(BD_MAX_NEWTON, std::max(s.P.newton_iterations, min_it));
    const int n_max = std::min(BD_MAX_NEWTON, std::max(std::max(s.P.newton_max, n_it), max_it));
    s.stats.newton_iterations = substep_body(s, n_it, n_max);
    return BD_OK;
  });
}

int bd_graph_capture(bd_solver* h, int pcg_iterations) {
  return guarded(h, [&](Solver& s) {
    if (pcg_iterations < 1) return fail(h, BD_ERR_INVALID, "bd_graph_capture: pcg_iterations must be positive");
    if (s.prof_on) return fail(h, BD_ERR_STATE, "bd_graph_capture: profiling is on (its events are host work)");
    if (s.npieces > 0) return fail(h, BD_ERR_STATE, "bd_graph_capture: detached pieces are updated by the host");
    if (s.learned != BD_LEARNED_OFF || learn_recording(s))
      return fail(h, BD_ERR_STATE, "bd_graph_capture: the learned path and sample recording run eagerly");
    const int n_it = std::min(BD_MAX_NEWTON, s.P.newton_iterations);
    if (std::max(s.P.newton_max, n_it) > n_it)
      return fail(h, BD_ERR_STATE, "bd_graph_capture: optional Newton iterations (newton_max > newton_iterations) are gated by the host");
    BD_HIP(hipStreamSynchronize(s.stream));
    release_graph(s);
    s.graph_pcg = pcg_iterations;
    BD_HIP(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
    try {
      substep_body(s, n_it, n_it);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s.stream, &g);
      if (g) (void)hipGraphDestroy(g);
      s.graph_pcg = 0;
      throw;
    }
    s.graph_pcg = 0;
    BD_HIP(hipStreamEndCapture(s.stream, &s.graph));
    BD_HIP(hipGraphInstantiate(&s.graph_exec, s.graph, nullptr, nullptr, 0));
    s.graph_newton = n_it;
    s.graph_gen = s.config_gen;
    s.graph_serial++;
    return BD_OK;
  });
}

int bd_graph_substep(bd_solver* h) {
  return guarded(h, [&](Solver& s) {
    if (!s.graph_exec) return fail(h, BD_ERR_STATE, "bd_graph_substep: no captured substep (bd_graph_capture)");
    if (s.graph_gen != s.config_gen)
      return fail(h, BD_ERR_STATE, "bd_graph_substep: the solver's configuration changed since bd_graph_capture; capture again");
    BD_HIP(hipGraphLaunch(s.graph_exec, s.stream));
    s.stats.newton_iterations = s.graph_newton;
    return BD_OK;
  });
}

int bd_graph_release(bd_solver* h) {
  return guarded(h, [&](Solver& s) {
    BD_HIP(hipStreamSynchronize(s.stream));
    release_graph(s);
    return BD_OK;
  });
}

int bd_profile_enable(bd_solver* h, int on) {
  return guarded(h, [&](Solver& s) {
    BD_HIP(hipStreamSynchronize(s.stream));
    s.prof_on = on != 0;
    s.config_gen++;
    s.prof_recs.clear();
    return BD_OK;
  });
}

int bd_profile_read(bd_solver* h, bd_kernel_profile* out, int max_entries, int* count) {
  return guarded(h, [&](Solver& s) {
    if (!out || !count) return fail(h, BD_ERR_INVALID, "bd_profile_read: null output");
    BD_HIP(hipStreamSynchronize(s.stream));
    bd_kernel_profile acc[PC_COUNT];
    std::memset(acc, 0, sizeof(acc));
    for (const ProfRec& r : s.prof_recs) {
      float ms = 0.f;
      BD_HIP(hipEventElapsedTime(&ms, s.prof_pool[r.ev], s.prof_pool[r.ev + 1]));
      acc[r.cls].launches += 1;
      acc[r.cls].total_ms += ms;
      acc[r.cls].bytes += r.bytes;
    }
    s.prof_recs.clear();
    int n = 0;
    for (int c = 0; c < PC_COUNT && n < max_entries; c++) {
      if (acc[c].launches == 0) continue;
      out[n] = acc[c];
      std::snprintf(out[n].name, sizeof(out[n].name), "%s", prof_name(c));
      n++;
    }
    *count = n;
    return BD_OK;
  });
}

int bd_learn_record(bd_solver* h, int capacity) {
  return guarded(h, [&](Solver& s) {
    if (capacity < 0) return fail(h, BD_ERR_INVALID, "bd_learn_record: negative capacity");
    BD_HIP(hipStreamSynchronize(s.stream));
    learn_record_start(s, capacity);
    s.config_gen++;
    return BD_OK;
  });
}

int bd_learn_train(bd_solver* h, int epochs, double lr, int batch, int targets, double* losses) {
  return guarded(h, [&](Solver& s) {
    if (epochs < 0 || !(lr > 0) || batch < 1) return fail(h, BD_ERR_INVALID, "bd_learn_train: epochs >= 0, lr > 0, batch >= 1");
    if (targets < 1 || targets > 3) return fail(h, BD_ERR_INVALID, "bd_learn_train: targets is 1 (warm start), 2 (slow modes) or 3");
    if (learn_sample_count(s) == 0) return fail(h, BD_ERR_STATE, "bd_learn_train: no recorded samples");
    learn_train(s, epochs, lr, batch, targets, losses);
    return BD_OK;
  });
}

int bd_set_learned(bd_solver* h, int on) {
  return guarded(h, [&](Solver& s) {
    if (on < BD_LEARNED_OFF || on > BD_LEARNED_DEFLATE) return fail(h, BD_ERR_INVALID, "bd_set_learned: unknown mode");
    if (on) {
      if (s.nlev < 2) return fail(h, BD_ERR_STATE, "bd_set_learned: the hierarchy has a single level");
      learn_alloc(s);
    }
    s.learned = on;
    s.config_gen++;
    // the next Newton iteration predicts from its hierarchy
    if (s.learned && s.hierarchy) learn_predict(s);
    return BD_OK;
  });
}

int bd_learn_get_weights(bd_solver* h, float* out, int count) {
  return guarded(h, [&](Solver& s) {
    if (!out || count != BD_NET_PARAMS) return fail(h, BD_ERR_INVALID, "bd_learn_get_weights: need BD_NET_PARAMS floats");
    if (s.nlev < 2) return fail(h, BD_ERR_STATE, "bd_learn_get_weights: the hierarchy has a single level");
    learn_alloc(s);
    learn_get_weights(s, out);
    return BD_OK;
  });
}

int bd_learn_set_weights(bd_solver* h, const float* w, int count) {
  return guarded(h, [&](Solver& s) {
    if (!w || count != BD_NET_PARAMS) return fail(h, BD_ERR_INVALID, "bd_learn_set_weights: need BD_NET_PARAMS floats");
    if (s.nlev < 2) return fail(h, BD_ERR_STATE, "bd_learn_set_weights: the hierarchy has a single level");
    learn_alloc(s);
    learn_set_weights(s, w);
    if (s.learned && s.hierarchy) learn_predict(s);
    return BD_OK;
  });
}

int bd_learn_loss_grad(bd_solver* h, int sample, double* loss, float* grad, int count) {
  return guarded(h, [&](Solver& s) {
    if (!loss || (grad && count != BD_NET_PARAMS)) return fail(h, BD_ERR_INVALID, "bd_learn_loss_grad: bad output");
    if (sample < 0 || sample >= learn_sample_count(s)) return fail(h, BD_ERR_INVALID, "bd_learn_loss_grad: no such sample");
    learn_loss_grad(s, sample, loss, grad);
    if (s.learned && s.hierarchy) learn_predict(s);
    return BD_OK;
  });
}

int bd_learn_read_sample(bd_solver* h, int sample, float* b, double* energy, float* op, uint8_t* fixed) {
  return guarded(h, [&](Solver& s) {
    if (sample < 0 || sample >= learn_sample_count(s)) return fail(h, BD_ERR_INVALID, "bd_learn_read_sample: no such sample");
    learn_read_sample(s, sample, b, energy, op, fixed);
    return BD_OK;
  });
}

int bd_learn_info(bd_solver* h, int* level, int* nodes, int* padded, int* samples) {
  return guarded(h, [&](Solver& s) {
    if (s.nlev < 2) return fail(h, BD_ERR_STATE, "bd_learn_info: the hierarchy has a single level");
    learn_alloc(s);
    if (level) *level = learn_level(s);
    if (nodes) *nodes = learn_nodes(s);
    if (padded) *padded = learn_padded(s);
    if (samples) *samples = learn_sample_count(s);
    return BD_OK;
  });
}

int bd_learn_read_fields(bd_solver* h, float* z) {
  return guarded(h, [&](Solver& s) {
    if (!z) return fail(h, BD_ERR_INVALID, "bd_learn_read_fields: null output");
    if (!learn_ready(s)) return fail(h, BD_ERR_STATE, "bd_learn_read_fields: no prediction yet");
    learn_read_z(s, z);
    return BD_OK;
  });
}

int bd_apply(bd_solver* h, int level, const float* x, float* y, double* xax, void* stream) {
  return guarded(h, [&](Solver& s) {
    if (level < 0 || level >= s.nlev) return fail(h, BD_ERR_INVALID, "bd_apply: no such level");
    if (!x || !y || !xax) return fail(h, BD_ERR_INVALID, "bd_apply: null pointer");
    if (!s.assembled) return fail(h, BD_ERR_STATE, "bd_apply: no assembled operator");
    if (level > 0 && !s.hierarchy) return fail(h, BD_ERR_STATE, "bd_apply: no Galerkin hierarchy");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s.stream;
    // a workspace per call: the caller's stream may run beside the solver's own work
    void* ws = nullptr;
    BD_HIP(hipMalloc(&ws, kApplyScratch));
    if (st != s.stream) BD_HIP(hipStreamSynchronize(s.stream));
    launch_apply_dot(s, level, x, y, static_cast<double*>(ws), xax, st);
    BD_HIP(hipGetLastError());
    BD_HIP(hipStreamSynchronize(st));
    BD_HIP(hipFree(ws));
    return BD_OK;
  });
}

int bd_apply_noalloc(bd_solver* h, int level, const float* x, float* y, double* xax, void* stream) {
  return guarded(h, [&](Solver& s) {
    if (level < 0 || level >= s.nlev) return fail(h, BD_ERR_INVALID, "bd_apply_noalloc: no such level");
    if (!x || !y || !xax) return fail(h, BD_ERR_INVALID, "bd_apply_noalloc: null pointer");
    if (!s.assembled) return fail(h, BD_ERR_STATE, "bd_apply_noalloc: no assembled operator");
    if (level > 0 && !s.hierarchy) return fail(h, BD_ERR_STATE, "bd_apply_noalloc: no Galerkin hierarchy");
    double* sp = nullptr;
    const hipError_t e = resolve_scratch(&sp);
    if (e != hipSuccess || !sp) return fail(h, BD_ERR_HIP, std::string("bd_apply_noalloc: scratch: ") + hipGetErrorString(e));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s.stream;
    if (st != s.stream) BD_HIP(hipStreamSynchronize(s.stream));
    launch_apply_dot(s, level, x, y, sp, xax, st);
    BD_HIP(hipGetLastError());
    BD_HIP(hipStreamSynchronize(st));
    return BD_OK;
  });
}

int bd_sync(bd_solver* h) {
  return guarded(h, [&](Solver& s) {
    coarse_join(s);
    BD_HIP(hipStreamSynchronize(s.stream));
    return BD_OK;
  });
}

size_t bd_buffer_bytes(const bd_solver* h, bd_buffer buf, int level) {
  if (!h) return 0;
  const Solver& s = h->s;
  int Nl = (level >= 0 && level < s.nlev) ? s.lv[level].N : 0;
  switch (buf) {
    case BD_BUF_X: case BD_BUF_V: case BD_BUF_DX: case BD_BUF_PRED: case BD_BUF_DELTA: case BD_BUF_RHS:
      return 3 * (size_t)s.N * sizeof(double);
    case BD_BUF_LEVEL_X: case BD_BUF_LEVEL_B: case BD_BUF_LEVEL_R:
      return 3 * (size_t)Nl * sizeof(double);
    case BD_BUF_STENCIL: return (size_t)Nl * 25 * 9 * sizeof(double);
    case BD_BUF_FIXED: return (size_t)Nl;
    case BD_BUF_PROLONG: return (level + 1 < s.nlev) ? 4 * (size_t)Nl * sizeof(double) : 0;
    case BD_BUF_REST: case BD_BUF_INTACT: case BD_BUF_FAILSITE: return 2 * (size_t)s.N * sizeof(double);
    case BD_BUF_CUTMASK: return 2 * (size_t)s.N * sizeof(uint32_t);
    case BD_BUF_SEGFLAGS: return 2 * (size_t)s.N * sizeof(uint16_t);
    case BD_BUF_TAILS: return 4 * (size_t)s.N * sizeof(double);
    case BD_BUF_FIELD: return 2 * (size_t)s.C;
    case BD_BUF_STATE: return (size_t)s.N;
    case BD_BUF_RIGIDK: return s.rigidk ? 12 * (size_t)s.N * sizeof(float) : 0;
    case BD_BUF_PIECES: return s.piece_frame && s.npieces > 0 ? 48 * (size_t)s.npieces * sizeof(double) : 0;
    case BD_BUF_CURVE: return s.curve ? 7 * (size_t)s.N * sizeof(float) : 0;
    case BD_BUF_LAST_UPDATE: return (size_t)s.N * sizeof(double);
    case BD_BUF_ENERGY: return 9 * sizeof(double);
    case BD_BUF_EXT: return 9 * (size_t)s.N * sizeof(float);
    case BD_BUF_CMETA: return 8 * (size_t)s.N * sizeof(float);
    case BD_BUF_PAIRS:
      return s.cells ? (kContactK + 2) * (size_t)s.N * sizeof(uint32_t) : (s.pcells ? kContactK * (size_t)s.N * sizeof(uint32_t) : 0);
    case BD_BUF_CELLS:
      return s.cells ? 2 * (size_t)kTable * sizeof(uint32_t) : (s.pcells ? 2 * (size_t)kPlyTable * sizeof(uint32_t) : 0);
    case BD_BUF_ITEMS: return s.cells ? 16 * (size_t)s.N * sizeof(float) : (s.pcells ? (size_t)s.N * sizeof(uint32_t) : 0);
    default: return 0;
  }
}

// expand a level operator to N x 25 x 9 (row-major 3x3 per stencil offset (dj+2)*5 + di+2)
static void expand_stencil(const Solver& s, int k, const std::vector<float>& op, double* out) {
  const Level& L = s.lv[k];
  const int Np = L.Np;
  std::memset(out, 0, (size_t)L.N * 225 * sizeof(double));
  auto put = [&](int n, int di, int dj, const float* b6) {
    double* d = out + ((size_t)n * 25 + (dj + 2) * 5 + (di + 2)) * 9;
    const float full[9] = {b6[0], b6[1], b6[2], b6[1], b6[3], b6[4], b6[2], b6[4], b6[5]};
    for (int q = 0; q < 9; q++) d[q] = full[q];
  };
  auto blk = [&](int base, int row, float* o) {
    for (int q = 0; q < 6; q++) o[q] = op[(size_t)(base + q) * Np + row];
  };
  for (int l = 0; l < L.geo.nl; l++) {
    int nx = L.geo.nx[l], ny = L.geo.ny[l], off = L.geo.off[l];
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int n = off + j * nx + i;
        float b[6];
        blk(0, n, b);
        put(n, 0, 0, b);
        if (L.fine) {
          const int fwd[4][3] = {{F_E, 1, 0}, {F_N, 0, 1}, {F_NE, 1, 1}, {F_NW, -1, 1}};
          for (auto& f : fwd) {
            int di = f[1], dj = f[2];
            if (i + di >= 0 && i + di < nx && j + dj < ny) {
              blk(f[0], n, b);
              put(n, di, dj, b);
            }
            if (i - di >= 0 && i - di < nx && j - dj >= 0) {
              blk(f[0], n - dj * nx - di, b);
              put(n, -di, -dj, b);
            }
          }
          const int sc[2][3] = {{F_EE, 2, 0}, {F_NN, 0, 2}};
          for (auto& f : sc) {
            int di = f[1], dj = f[2];
            float v;
            if (i + di < nx && j + dj < ny) {
              v = op[(size_t)f[0] * Np + n];
              float s6[6] = {v, 0, 0, v, 0, v};
              put(n, di, dj, s6);
            }
            if (i - di >= 0 && j - dj >= 0) {
              v = op[(size_t)f[0] * Np + n - dj * nx - di];
              float s6[6] = {v, 0, 0, v, 0, v};
              put(n, -di, -dj, s6);
            }
          }
        } else {
          for (int kb = 1; kb < kCoarseBlocks; kb++) {
            int di = coarse_di(kb), dj = coarse_dj(kb);
            if (i + di >= 0 && i + di < nx && j + dj < ny) {
              blk(6 * kb, n, b);
              put(n, di, dj, b);
            }
            if (i - di >= 0 && i - di < nx && j - dj >= 0) {
              blk(6 * kb, n - dj * nx - di, b);
              put(n, -di, -dj, b);
            }
          }
        }
      }
  }
}

int bd_read(bd_solver* h, bd_buffer buf, int level, void* dst, size_t bytes) {
  return guarded(h, [&](Solver& s) {
    size_t need = bd_buffer_bytes(h, buf, level);
    if (need == 0) return fail(h, BD_ERR_INVALID, "bd_read: unknown buffer or level");
    if (!dst || bytes < need) return fail(h, BD_ERR_INVALID, "bd_read: destination too small");
    double* d = static_cast<double*>(dst);
    switch (buf) {
      case BD_BUF_X: download_vec3(s, s.x, s.N, s.Np, d); break;
      case BD_BUF_V: download_vec3(s, s.v, s.N, s.Np, d); break;
      case BD_BUF_DX: download_vec3(s, s.dx, s.N, s.Np, d); break;
      case BD_BUF_PRED: download_vec3(s, s.pred, s.N, s.Np, d); break;
      case BD_BUF_DELTA: {
        std::vector<double> t(3 * (size_t)s.Np);
        download(s, t.data(), s.delta, t.size());
        for (int n = 0; n < s.N; n++)
          for (int q = 0; q < 3; q++) d[3 * n + q] = t[(size_t)q * s.Np + n];
        break;
      }
      case BD_BUF_RHS: download_vec3(s, s.rhs, s.N, s.Np, d); break;
      case BD_BUF_LEVEL_X: download_vec3(s, s.lv[level].x, s.lv[level].N, s.lv[level].Np, d); break;
      case BD_BUF_LEVEL_B: download_vec3(s, s.lv[level].b, s.lv[level].N, s.lv[level].Np, d); break;
      case BD_BUF_LEVEL_R: download_vec3(s, s.lv[level].r, s.lv[level].N, s.lv[level].Np, d); break;
      case BD_BUF_STENCIL: {
        const Level& L = s.lv[level];
        std::vector<float> op((size_t)(L.fine ? kFineComps : kCoarseComps) * L.Np);
        download(s, op.data(), L.op, op.size());
        expand_stencil(s, level, op, d);
        break;
      }
      case BD_BUF_FIXED: download(s, static_cast<uint8_t*>(dst), s.lv[level].fixed, (size_t)s.lv[level].N); break;
      case BD_BUF_PROLONG: {
        const Level& L = s.lv[level];
        std::vector<float> pw(4 * (size_t)L.Np);
        download(s, pw.data(), L.pw, pw.size());
        for (int n = 0; n < L.N; n++)
          for (int q = 0; q < 4; q++) d[4 * n + q] = pw[(size_t)q * L.Np + n];
        break;
      }
      case BD_BUF_REST: download(s, d, s.rest, 2 * (size_t)s.N); break;
      case BD_BUF_INTACT: download(s, d, s.intact, 2 * (size_t)s.N); break;
      case BD_BUF_FAILSITE: download(s, d, s.fail_site, 2 * (size_t)s.N); break;
      case BD_BUF_TAILS: download(s, d, s.tails, 4 * (size_t)s.N); break;
      case BD_BUF_CUTMASK: download(s, static_cast<uint32_t*>(dst), s.cut_mask, 2 * (size_t)s.N); break;
      case BD_BUF_SEGFLAGS: download(s, static_cast<uint16_t*>(dst), s.flags, 2 * (size_t)s.N); break;
      case BD_BUF_FIELD: download(s, static_cast<uint8_t*>(dst), s.damage, 2 * (size_t)s.C); break;
      case BD_BUF_STATE: download(s, static_cast<uint8_t*>(dst), s.nstate, (size_t)s.N); break;
      case BD_BUF_RIGIDK:
      case BD_BUF_CURVE: {
        const int planes = buf == BD_BUF_RIGIDK ? 12 : 7;
        std::vector<float> t((size_t)planes * s.Np);
        download(s, t.data(), buf == BD_BUF_RIGIDK ? s.rigidk : s.curve, t.size());
        float* o = static_cast<float*>(dst);
        for (int n = 0; n < s.N; n++)
          for (int q = 0; q < planes; q++) o[(size_t)planes * n + q] = t[(size_t)q * s.Np + n];
        break;
      }
      case BD_BUF_PIECES: {
        std::vector<double> fr((size_t)s.npieces * kPieceFrame), st((size_t)s.npieces * 8);
        download(s, fr.data(), s.piece_frame, fr.size());
        download(s, st.data(), s.piece_step, st.size());
        double* o = static_cast<double*>(dst);
        for (int p = 0; p < s.npieces; p++) {
          for (int q = 0; q < kPieceFrame; q++) o[48 * (size_t)p + q] = fr[(size_t)p * kPieceFrame + q];
          for (int q = 0; q < 8; q++) o[48 * (size_t)p + kPieceFrame + q] = st[(size_t)p * 8 + q];
        }
        break;
      }
      case BD_BUF_LAST_UPDATE: {
        std::vector<float> t(s.N);
        download(s, t.data(), s.last_update, (size_t)s.N);
        for (int n = 0; n < s.N; n++) d[n] = t[n];
        break;
      }
      case BD_BUF_ENERGY: {
        download(s, s.scal_h, s.scal, 1);
        for (int k = 0; k < 4; k++) {
          d[k] = s.scal_h->e[k];
          d[4 + k] = s.scal_h->m[k];
        }
        d[8] = s.scal_h->line_alpha;
        break;
      }
      case BD_BUF_EXT: {
        std::vector<float> t(9 * (size_t)s.Np);
        download(s, t.data(), s.ext, t.size());
        float* o = static_cast<float*>(dst);
        for (int n = 0; n < s.N; n++)
          for (int q = 0; q < 9; q++) o[9 * (size_t)n + q] = t[(size_t)q * s.Np + n];
        break;
      }
      case BD_BUF_CMETA: download(s, static_cast<float4*>(dst), s.cmeta, 2 * (size_t)s.N); break;
      case BD_BUF_PAIRS: {
        std::vector<uint32_t> t(kContactK * (size_t)s.Np);
        if (s.contact_mode == kContactPlies) {
          download(s, t.data(), s.pairs, t.size());
          uint32_t* o = static_cast<uint32_t*>(dst);
          for (int n = 0; n < s.N; n++)
            for (int q = 0; q < kContactK; q++) o[kContactK * (size_t)n + q] = t[(size_t)q * s.Np + n];
          break;
        }
        std::vector<unsigned long long> k(s.N);
        std::vector<uint32_t> m(s.N);
        download(s, t.data(), s.pairs, t.size());
        download(s, k.data(), s.keys, (size_t)s.N);
        download(s, m.data(), s.marks, (size_t)s.N);
        uint32_t* o = static_cast<uint32_t*>(dst);
        for (int n = 0; n < s.N; n++) {
          for (int q = 0; q < kContactK; q++) o[(kContactK + 2) * (size_t)n + q] = t[(size_t)q * s.Np + n];
          o[(kContactK + 2) * (size_t)n + kContactK] = (uint32_t)(k[n] & 0xFFFFFFFFull);  // node + 1 of a pending key
          o[(kContactK + 2) * (size_t)n + kContactK + 1] = m[n];
        }
        break;
      }
      case BD_BUF_CELLS:
        if (s.contact_mode == kContactPlies) download(s, static_cast<uint32_t*>(dst), s.pcells, 2 * (size_t)kPlyTable);
        else download(s, static_cast<uint32_t*>(dst), s.cells, 2 * (size_t)kTable);
        break;
      case BD_BUF_ITEMS:
        if (s.contact_mode == kContactPlies) download(s, static_cast<uint32_t*>(dst), s.pitems, (size_t)s.N);
        else download(s, static_cast<float4*>(dst), s.items, 4 * (size_t)s.N);
        break;
      default: return fail(h, BD_ERR_INVALID, "bd_read: unsupported buffer");
    }
    return BD_OK;
  });
}

int bd_write(bd_solver* h, bd_buffer buf, int level, const void* src, size_t bytes) {
  return guarded(h, [&](Solver& s) {
    size_t need = bd_buffer_bytes(h, buf, level);
    if (need == 0) return fail(h, BD_ERR_INVALID, "bd_write: unknown buffer or level");
    if (!src || bytes < need) return fail(h, BD_ERR_INVALID, "bd_write: source too small");
    const double* d = static_cast<const double*>(src);
    switch (buf) {
      case BD_BUF_X: upload_vec3(s, s.x, s.N, s.Np, d); break;
      case BD_BUF_V: upload_vec3(s, s.v, s.N, s.Np, d); break;
      case BD_BUF_DX: upload_vec3(s, s.dx, s.N, s.Np, d); break;
      case BD_BUF_PRED: upload_vec3(s, s.pred, s.N, s.Np, d); break;
      case BD_BUF_DELTA: {
        std::vector<double> t(3 * (size_t)s.Np, 0.0);
        for (int n = 0; n < s.N; n++)
          for (int q = 0; q < 3; q++) t[(size_t)q * s.Np + n] = d[3 * n + q];
        upload(s, s.delta, t.data(), t.size());
        break;
      }
      case BD_BUF_RHS: upload_vec3(s, s.rhs, s.N, s.Np, d); break;
      case BD_BUF_LEVEL_X:
        upload_vec3(s, s.lv[level].x, s.lv[level].N, s.lv[level].Np, d);
        launch_colmajor_x(s, level);
        break;
      case BD_BUF_LEVEL_B: upload_vec3(s, s.lv[level].b, s.lv[level].N, s.lv[level].Np, d); break;
      case BD_BUF_REST: upload(s, s.rest, d, 2 * (size_t)s.N); break;
      case BD_BUF_INTACT: upload(s, s.intact, d, 2 * (size_t)s.N); break;
      case BD_BUF_FAILSITE: upload(s, s.fail_site, d, 2 * (size_t)s.N); break;
      case BD_BUF_TAILS: upload(s, s.tails, d, 4 * (size_t)s.N); break;
      case BD_BUF_CUTMASK: upload(s, s.cut_mask, static_cast<const uint32_t*>(src), 2 * (size_t)s.N); break;
      case BD_BUF_SEGFLAGS: upload(s, s.flags, static_cast<const uint16_t*>(src), 2 * (size_t)s.N); break;
      case BD_BUF_FIELD: upload(s, s.damage, static_cast<const uint8_t*>(src), 2 * (size_t)s.C); break;
      case BD_BUF_STATE: {
        const uint8_t* in = static_cast<const uint8_t*>(src);
        for (int n = 0; n < s.N; n++)
          if (in[n] > kStateRetired) return fail(h, BD_ERR_INVALID, "bd_write: node state must be 0, 1 or 2");
        for (int n = 0; n < s.N; n++) s.any_detached = s.any_detached || in[n] != kStateCloth;
        s.config_gen++;
        upload(s, s.nstate, in, (size_t)s.N);
        launch_apply_state(s);
        break;
      }
      case BD_BUF_EXT: {
        const float* in = static_cast<const float*>(src);
        std::vector<float> t(9 * (size_t)s.Np, 0.f);
        for (int n = 0; n < s.N; n++)
          for (int q = 0; q < 9; q++) t[(size_t)q * s.Np + n] = in[9 * (size_t)n + q];
        upload(s, s.ext, t.data(), t.size());
        break;
      }
      case BD_BUF_CMETA: upload(s, s.cmeta, static_cast<const float4*>(src), 2 * (size_t)s.N); break;
      default: return fail(h, BD_ERR_INVALID, "bd_write: buffer is read-only");
    }
    return BD_OK;
  });
}

int bd_get_stats(bd_solver* h, bd_stats* out) {
  return guarded(h, [&](Solver& s) {
    if (!out) return fail(h, BD_ERR_INVALID, "bd_get_stats: null output");
    download(s, s.scal_h, s.scal, 1);
    int counters[8];
    download(s, counters, s.counters, 8);
    bd_stats st = s.stats;
    for (int k = 0; k < BD_MAX_NEWTON; k++) {
      st.pcg_iterations[k] = s.scal_h->rec_iters[k];
      st.rel_residual[k] = s.scal_h->rec_relres[k];
      st.alpha[k] = s.scal_h->rec_alpha[k];
      float mu;
      std::memcpy(&mu, &s.scal_h->max_update[k], sizeof(float));
      st.max_update[k] = mu;
    }
    st.rejections = s.scal_h->rejections;
    st.poll_fallbacks = s.poll_fallbacks;
    st.learned_fallbacks = s.scal_h->learn_fallbacks;
    for (int k = 0; k < BD_MAX_NEWTON; k++) st.warm_relres[k] = s.scal_h->rec_warm[k];
    st.ruptures = counters[0];
    st.pulled_out = counters[1];
    float vmax;
    std::memcpy(&vmax, &counters[2], sizeof(float));
    st.max_speed = vmax;
    // the largest speed counts from this read on (rest detection reads and clears it)
    BD_HIP(hipMemsetAsync(s.counters + 2, 0, sizeof(int), s.stream));
    *out = st;
    return BD_OK;
  });
}

}  // extern "C"

// ------------------------------------------------------------------ batches of independent solvers

struct bd_batch {
  std::vector<bd_solver*> solvers;
  // concurrent form: each solver's eager substep on its own stream, `threads` host threads taking
  // the solvers in `order` (largest system first)
  bool concurrent = false;
  int threads = 0;
  int device = 0;
  std::vector<int> order;
  // fused form: every solver in BD_PCG_WORKGROUP mode, the PCG solves of a Newton iteration in one
  // launch over the systems' descriptors (copied back to back)
  bool fused = false;
  void* systems = nullptr;
  size_t lds = 0;
  int newton = 0;
  std::vector<unsigned> gens;
  std::vector<unsigned> captured;    // each solver's capture serial when the batch was made
  std::vector<hipEvent_t> ready;     // per solver: its stream's work before a launch
  hipEvent_t done = nullptr;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string err;
};

namespace {

void batch_free(bd_batch* b) {
  if (b->systems) (void)hipFree(b->systems);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (hipEvent_t e : b->ready) (void)hipEventDestroy(e);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

template <class F>
int batch_guarded(bd_batch* b, F&& f) {
  if (!b) return BD_ERR_INVALID;
  try {
    return f(*b);
  } catch (const std::exception& e) {
    b->err = e.what();
    return BD_ERR_HIP;
  }
}

}  // namespace

extern "C" {

const char* bd_batch_last_error(const bd_batch* b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int bd_batch_create(bd_solver* const* solvers, int count, bd_batch** out) {
  g_create_error.clear();
  if (!out || !solvers || count < 1) {
    g_create_error = "bd_batch_create: need at least one solver and an output";
    return BD_ERR_INVALID;
  }
  *out = nullptr;
  bool fused = true;
  for (int i = 0; i < count; i++) {
    if (!solvers[i]) {
      g_create_error = "bd_batch_create: every solver needs a captured substep (bd_graph_capture) or BD_PCG_WORKGROUP";
      return BD_ERR_STATE;
    }
    for (int j = 0; j < i; j++)
      if (solvers[j] == solvers[i]) {
        g_create_error = "bd_batch_create: a solver appears twice";
        return BD_ERR_INVALID;
      }
    fused = fused && solvers[i]->s.pcg_mode == BD_PCG_WORKGROUP;
  }
  const int newton = fused ? std::min(BD_MAX_NEWTON, solvers[0]->s.P.newton_iterations) : 0;
  for (int i = 0; i < count; i++) {
    const Solver& s = solvers[i]->s;
    if (fused) {
      if (std::min(BD_MAX_NEWTON, s.P.newton_iterations) != newton || std::max(s.P.newton_max, s.P.newton_iterations) > s.P.newton_iterations) {
        g_create_error = "bd_batch_create: solvers in lockstep need the same Newton iterations and no optional ones";
        return BD_ERR_STATE;
      }
    } else if (!s.graph_exec) {
      g_create_error = "bd_batch_create: every solver needs a captured substep (bd_graph_capture) or BD_PCG_WORKGROUP";
      return BD_ERR_STATE;
    }
  }
  bd_batch* b = new bd_batch;
  try {
    BD_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    BD_HIP(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    b->fused = fused;
    b->newton = newton;
    if (fused) {
      const size_t one = wg_sys_bytes();
      BD_HIP(hipMalloc(&b->systems, one * count));
      for (int i = 0; i < count; i++) {
        Solver& s = solvers[i]->s;
        BD_HIP(hipMemcpy(static_cast<char*>(b->systems) + one * i, s.wg_sys, one, hipMemcpyDeviceToDevice));
        b->lds = std::max(b->lds, s.wg_lds);
      }
    } else {
      BD_HIP(hipGraphCreate(&b->graph, 0));
    }
    for (int i = 0; i < count; i++) {
      Solver& s = solvers[i]->s;
      if (!fused) {
        // each patch its own branch: no dependencies between them
        hipGraphNode_t node;
        BD_HIP(hipGraphAddChildGraphNode(&node, b->graph, nullptr, 0, s.graph));
      }
      hipEvent_t e;
      BD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      b->ready.push_back(e);
      b->solvers.push_back(solvers[i]);
      b->captured.push_back(s.graph_serial);
      b->gens.push_back(s.config_gen);
    }
    if (!fused) BD_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  } catch (const std::exception& e) {
    g_create_error = e.what();
    batch_free(b);
    return BD_ERR_HIP;
  }
  *out = b;
  return BD_OK;
}

int bd_batch_create_concurrent(bd_solver* const* solvers, int count, int threads, bd_batch** out) {
  g_create_error.clear();
  if (!out || !solvers || count < 1 || threads < 0) {
    g_create_error = "bd_batch_create_concurrent: need at least one solver, an output and threads >= 0";
    return BD_ERR_INVALID;
  }
  *out = nullptr;
  for (int i = 0; i < count; i++) {
    if (!solvers[i]) {
      g_create_error = "bd_batch_create_concurrent: null solver";
      return BD_ERR_INVALID;
    }
    for (int j = 0; j < i; j++)
      if (solvers[j] == solvers[i]) {
        g_create_error = "bd_batch_create_concurrent: a solver appears twice";
        return BD_ERR_INVALID;
      }
  }
  bd_batch* b = new bd_batch;
  b->concurrent = true;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  b->threads = std::min(count, threads > 0 ? threads : (int)std::min(8u, hw));
  if (hipGetDevice(&b->device) != hipSuccess) {
    g_create_error = "bd_batch_create_concurrent: no current device";
    delete b;
    return BD_ERR_HIP;
  }
  b->solvers.assign(solvers, solvers + count);
  // largest systems first, so that the last ones to start are the shortest
  b->order.resize(count);
  for (int i = 0; i < count; i++) b->order[i] = i;
  std::stable_sort(b->order.begin(), b->order.end(),
                   [&](int u, int v) { return b->solvers[u]->s.N > b->solvers[v]->s.N; });
  *out = b;
  return BD_OK;
}

int bd_batch_substep(bd_batch* b) {
  return batch_guarded(b, [&](bd_batch& B) {
    if (B.concurrent) {
      // each worker takes the next solver and runs its whole substep (the host's convergence checks
      // included) on the solver's own stream; the streams' kernels then share the chip
      std::atomic<int> next{0};
      std::mutex mu;
      int rc = BD_OK;
      std::string err;
      auto work = [&](bool own_thread) {
        if (own_thread && hipSetDevice(B.device) != hipSuccess) {
          std::lock_guard<std::mutex> lock(mu);
          if (rc == BD_OK) {
            rc = BD_ERR_HIP;
            err = "bd_batch_substep: hipSetDevice failed in a worker";
          }
          return;
        }
        for (;;) {
          const int k = next.fetch_add(1);
          if (k >= (int)B.order.size()) return;
          const int i = B.order[k];
          // a solver with a captured substep replays it, the others run it eagerly
          const int r = B.solvers[i]->s.graph_exec ? bd_graph_substep(B.solvers[i]) : bd_substep(B.solvers[i], 0, 0);
          if (r != BD_OK) {
            std::lock_guard<std::mutex> lock(mu);
            if (rc == BD_OK) {
              rc = r;
              err = "bd_batch_substep: solver " + std::to_string(i) + ": " + B.solvers[i]->s.err;
            }
          }
        }
      };
      std::vector<std::thread> pool;
      for (int t = 1; t < B.threads; t++) pool.emplace_back(work, true);
      work(false);
      for (std::thread& t : pool) t.join();
      if (rc != BD_OK) B.err = err;
      return rc;
    }
    for (size_t i = 0; i < B.solvers.size(); i++) {
      const Solver& s = B.solvers[i]->s;
      const bool stale = B.fused ? (s.config_gen != B.gens[i] || s.pcg_mode != BD_PCG_WORKGROUP)
                                 : (!s.graph_exec || s.graph_serial != B.captured[i] || s.graph_gen != s.config_gen);
      if (stale) {
        B.err = "bd_batch_substep: solver " + std::to_string(i) + " was captured again or changed since bd_batch_create";
        return BD_ERR_STATE;
      }
    }
    if (B.fused) {
      // the substeps in lockstep on their own streams; each Newton iteration's solves in one launch
      for (bd_solver* h : B.solvers) substep_begin(h->s);
      for (int it = 0; it < B.newton; it++) {
        for (size_t i = 0; i < B.solvers.size(); i++) {
          Solver& s = B.solvers[i]->s;
          s.newton_it = it;
          newton_pre(s, it == 0 || !s.P.lag_coarse);
          launch_pcg_setup(s);
          coarse_join(s);
          BD_HIP(hipEventRecord(B.ready[i], s.stream));
          BD_HIP(hipStreamWaitEvent(B.stream, B.ready[i], 0));
        }
        launch_pcg_wg_many(B.systems, (int)B.solvers.size(), B.lds, B.stream);
        BD_HIP(hipGetLastError());
        BD_HIP(hipEventRecord(B.done, B.stream));
        for (bd_solver* h : B.solvers) {
          BD_HIP(hipStreamWaitEvent(h->s.stream, B.done, 0));
          newton_post(h->s, it);
        }
      }
      for (bd_solver* h : B.solvers) {
        substep_end(h->s);
        h->s.stats.newton_iterations = B.newton;
      }
      return BD_OK;
    }
    // the launch follows each solver's queued work (grip goals, uploads), and each solver's later
    // work follows the launch
    for (size_t i = 0; i < B.solvers.size(); i++) {
      BD_HIP(hipEventRecord(B.ready[i], B.solvers[i]->s.stream));
      BD_HIP(hipStreamWaitEvent(B.stream, B.ready[i], 0));
    }
    BD_HIP(hipGraphLaunch(B.exec, B.stream));
    BD_HIP(hipEventRecord(B.done, B.stream));
    for (bd_solver* h : B.solvers) {
      BD_HIP(hipStreamWaitEvent(h->s.stream, B.done, 0));
      h->s.stats.newton_iterations = h->s.graph_newton;
    }
    return BD_OK;
  });
}

int bd_batch_sync(bd_batch* b) {
  return batch_guarded(b, [&](bd_batch& B) {
    if (B.stream) BD_HIP(hipStreamSynchronize(B.stream));
    for (bd_solver* h : B.solvers) BD_HIP(hipStreamSynchronize(h->s.stream));
    return BD_OK;
  });
}

int bd_batch_destroy(bd_batch* b) {
  if (!b) return BD_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_free(b);
  return BD_OK;
}

}  // extern "C"

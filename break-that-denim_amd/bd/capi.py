"""ctypes mirror of include/bd.h and the loader of the in-tree libbd.so.

`SceneBuffers` marshals a CompiledScene into a `bd_scene` struct and keeps the numpy arrays it
points at alive for as long as the struct is used.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent.parent          # break-that-denim_amd/
LIB_PATH = PKG_DIR / "libbd.so"

BD_OK = 0
BD_ERR_INVALID, BD_ERR_HIP, BD_ERR_LIMIT, BD_ERR_NOMEM, BD_ERR_STATE = -1, -2, -3, -4, -5
BD_PCG_GRID, BD_PCG_WORKGROUP = 0, 1
BD_MAX_LINE = 1024
BD_MAX_LEVELS = 12
BD_MAX_NEWTON = 8
BD_MAX_GRIPS = 16
BD_PRECISION_FP32 = 0
BD_PRECISION_FP16 = 1
BD_COARSE_SWEEPS = 0
BD_COARSE_EXACT = 1
BD_COARSE_EXACT_LAGGED = 2
BD_SMOOTHER_DEFAULT = -1
BD_SMOOTHER_DIRECT = 0
BD_SMOOTHER_STAGED4 = 4
BD_SMOOTHER_STAGED8 = 8
BD_SMOOTHER_SPLIT = 16
BD_SMOOTHER_SPLIT4 = 17
BD_SMOOTHER_FACTORED = 18

(BD_BUF_X, BD_BUF_V, BD_BUF_DX, BD_BUF_PRED, BD_BUF_DELTA, BD_BUF_RHS, BD_BUF_LEVEL_X, BD_BUF_LEVEL_B,
 BD_BUF_LEVEL_R, BD_BUF_STENCIL, BD_BUF_FIXED, BD_BUF_PROLONG, BD_BUF_REST, BD_BUF_INTACT, BD_BUF_CUTMASK,
 BD_BUF_SEGFLAGS, BD_BUF_FAILSITE, BD_BUF_TAILS, BD_BUF_FIELD, BD_BUF_LAST_UPDATE, BD_BUF_ENERGY,
 BD_BUF_EXT, BD_BUF_CMETA, BD_BUF_PAIRS, BD_BUF_CELLS, BD_BUF_ITEMS, BD_BUF_STATE, BD_BUF_RIGIDK, BD_BUF_PIECES,
 BD_BUF_CURVE) = range(30)

BD_MAX_PIECES = 4096
BD_NO_PIECE = 0xFFFFFFFF

BD_CONTACT_K = 4
BD_CONTACT_TABLE_BITS = 17
BD_CONTACT_NONE = 0xFFFFFFFF
BD_CONTACT_MAX_NODES = 1 << 24

BD_NET_FIELDS = 8
BD_NET_PARAMS = 37216
BD_NET_SLOW_ITER = 4
BD_LEARNED_OFF, BD_LEARNED_WARM, BD_LEARNED_DEFLATE = 0, 1, 2


class bd_yarn(C.Structure):
    _fields_ = [("linear_density", C.c_double), ("ea_low", C.c_double), ("ea_high", C.c_double),
                ("crimp_strain", C.c_double), ("ea_compress", C.c_double), ("axial_viscosity", C.c_double),
                ("bend_rigidity", C.c_double), ("bend_viscosity", C.c_double), ("strength", C.c_double),
                ("softening_opening", C.c_double), ("yarns_per_span", C.c_int32), ("pad_", C.c_int32)]


class bd_material(C.Structure):
    _fields_ = [("pitch_u", C.c_double), ("pitch_v", C.c_double), ("yarns", bd_yarn * 2),
                ("shear_stiffness", C.c_double), ("crossing_friction", C.c_double),
                ("crossing_normal0", C.c_double), ("crossing_wrap", C.c_double), ("air_drag", C.c_double),
                ("thickness", C.c_double)]


class bd_layer(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("node_offset", C.c_int32), ("bundle_u", C.c_int32),
                ("bundle_v", C.c_int32), ("field_offset", C.c_int32), ("field_k", C.c_int32),
                ("field_m", C.c_int32), ("weave_pitch_u", C.c_double), ("weave_pitch_v", C.c_double),
                ("material", bd_material)]


class bd_scene(C.Structure):
    _fields_ = [("node_count", C.c_int32), ("layer_count", C.c_int32), ("layers", C.POINTER(bd_layer)),
                ("position", C.POINTER(C.c_double)), ("mass", C.POINTER(C.c_double)),
                ("pinned", C.POINTER(C.c_uint8)), ("rest_length", C.POINTER(C.c_double)),
                ("softening", C.POINTER(C.c_double)), ("seg_flags", C.POINTER(C.c_uint16)),
                ("intact", C.POINTER(C.c_double)), ("cut_mask", C.POINTER(C.c_uint32)),
                ("fail_site", C.POINTER(C.c_double)), ("crossing_normal0", C.POINTER(C.c_double)),
                ("tails", C.POINTER(C.c_double)), ("crossing_count", C.c_int32), ("pad_", C.c_int32),
                ("field_strength", C.POINTER(C.c_float)), ("field_damage", C.POINTER(C.c_uint8)),
                ("field_area", C.POINTER(C.c_uint8))]


class bd_params(C.Structure):
    _fields_ = [("h", C.c_double), ("gravity", C.c_double * 3), ("newton_iterations", C.c_int32),
                ("newton_max", C.c_int32), ("newton_tolerance", C.c_double), ("pcg_iterations", C.c_int32),
                ("pcg_max_iterations", C.c_int32), ("pcg_rel_tol", C.c_double), ("coarse_sweeps", C.c_int32),
                ("max_coarsest_nodes", C.c_int32), ("lag_coarse", C.c_int32), ("line_search", C.c_int32),
                ("history", C.c_int32), ("precision", C.c_int32), ("self_contact", C.c_int32), ("coarse_solve", C.c_int32),
                ("smooth_sweeps", C.c_int32), ("net_level", C.c_int32)]


class bd_grips(C.Structure):
    _fields_ = [("grip_count", C.c_int32), ("entry_count", C.c_int32), ("node", C.POINTER(C.c_int32)),
                ("grip", C.POINTER(C.c_int32)), ("weight", C.POINTER(C.c_double)),
                ("offset", C.POINTER(C.c_double)), ("goal", C.POINTER(C.c_double)),
                ("stiffness", C.POINTER(C.c_double))]


class bd_stats(C.Structure):
    _fields_ = [("newton_iterations", C.c_int32), ("pcg_iterations", C.c_int32 * BD_MAX_NEWTON),
                ("rel_residual", C.c_double * BD_MAX_NEWTON), ("alpha", C.c_double * BD_MAX_NEWTON),
                ("max_update", C.c_double * BD_MAX_NEWTON), ("rejections", C.c_int32),
                ("ruptures", C.c_int32), ("pulled_out", C.c_int32), ("level_count", C.c_int32),
                ("level_nodes", C.c_int32 * BD_MAX_LEVELS), ("poll_fallbacks", C.c_int32),
                ("learned_fallbacks", C.c_int32), ("warm_relres", C.c_double * BD_MAX_NEWTON),
                ("max_speed", C.c_double)]


class bd_kernel_profile(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int32), ("pad_", C.c_int32), ("total_ms", C.c_double),
                ("bytes", C.c_double)]


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _yarn_struct(y) -> bd_yarn:
    return bd_yarn(y["linearDensity"], y["eaLow"], y["eaHigh"], y["crimpStrain"], y["eaCompress"],
                   y["axialViscosity"], y["bendRigidity"], y["bendViscosity"], y["strength"],
                   y["softeningOpening"], int(y.get("yarnsPerSpan", 1)), 0)


def layer_struct(L) -> bd_layer:
    m = L.material
    mat = bd_material(m["pitchU"], m["pitchV"], (bd_yarn * 2)(_yarn_struct(m["yarns"][0]), _yarn_struct(m["yarns"][1])),
                      m["shearStiffness"], m["crossingFriction"], m["crossingNormal0"], m["crossingWrap"], m["airDrag"],
                      m.get("thickness", 0.0))
    return bd_layer(L.nx, L.ny, L.node_offset, L.bu, L.bv, L.field_offset, L.K, L.M,
                    L.weave["pitchU"], L.weave["pitchV"], mat)


class SceneBuffers:
    """A `bd_scene` pointing into contiguous copies of a CompiledScene's arrays."""

    def __init__(self, S):
        self.N = S.node_count
        self.keep = {
            "position": np.ascontiguousarray(S.position, dtype=np.float64).ravel(),
            "mass": np.ascontiguousarray(S.mass, dtype=np.float64),
            "pinned": np.ascontiguousarray(S.pinned, dtype=np.uint8),
            "rest_length": np.ascontiguousarray(S.rest_length, dtype=np.float64),
            "softening": np.ascontiguousarray(S.softening, dtype=np.float64),
            "seg_flags": np.ascontiguousarray(S.seg_flags, dtype=np.uint16),
            "intact": np.ascontiguousarray(S.intact, dtype=np.float64),
            "cut_mask": np.ascontiguousarray(S.cut_mask, dtype=np.uint32),
            "fail_site": np.ascontiguousarray(S.fail_site, dtype=np.float64),
            "crossing_normal0": np.ascontiguousarray(S.crossing_normal0, dtype=np.float64),
            "tails": np.ascontiguousarray(S.tails, dtype=np.float64),
            "field_strength": np.ascontiguousarray(S.field_strength, dtype=np.float32),
            "field_damage": np.ascontiguousarray(S.field_damage, dtype=np.uint8),
            "field_area": np.ascontiguousarray(S.field_area, dtype=np.uint8),
        }
        self.layers = (bd_layer * len(S.layers))(*[layer_struct(L) for L in S.layers])
        k = self.keep
        self.scene = bd_scene(
            S.node_count, len(S.layers), C.cast(self.layers, C.POINTER(bd_layer)),
            _ptr(k["position"], C.c_double), _ptr(k["mass"], C.c_double), _ptr(k["pinned"], C.c_uint8),
            _ptr(k["rest_length"], C.c_double), _ptr(k["softening"], C.c_double),
            _ptr(k["seg_flags"], C.c_uint16), _ptr(k["intact"], C.c_double), _ptr(k["cut_mask"], C.c_uint32),
            _ptr(k["fail_site"], C.c_double), _ptr(k["crossing_normal0"], C.c_double),
            _ptr(k["tails"], C.c_double), S.crossing_count, 0, _ptr(k["field_strength"], C.c_float),
            _ptr(k["field_damage"], C.c_uint8), _ptr(k["field_area"], C.c_uint8))


def make_params(h=1 / 120, gravity=(0.0, 0.0, -9.81), newton_iterations=2, newton_max=None,
                newton_tolerance=1e-5, pcg_iterations=2, pcg_rel_tol=0.0, pcg_max_iterations=200,
                coarse_sweeps=2, max_coarsest_nodes=400, lag_coarse=True, line_search=True, history=True,
                precision=BD_PRECISION_FP32, self_contact=False, coarse_solve=BD_COARSE_SWEEPS,
                smooth_sweeps=1, net_level=0) -> bd_params:
    """Solver parameters; the defaults are the reference's production schedule (GPU_DEFAULTS,
    gpu/solver.ts:61-72: h 1/120, 2 Newton x 2 PCG, 2 coarse sweeps, coarsest <= 400 nodes)."""
    p = bd_params()
    p.h = h
    p.gravity[:] = list(gravity)
    p.newton_iterations = newton_iterations
    p.newton_max = newton_iterations if newton_max is None else newton_max
    p.newton_tolerance = newton_tolerance
    p.pcg_iterations = pcg_iterations
    p.pcg_max_iterations = pcg_max_iterations
    p.pcg_rel_tol = pcg_rel_tol
    p.coarse_sweeps = coarse_sweeps
    p.max_coarsest_nodes = max_coarsest_nodes
    p.lag_coarse = int(bool(lag_coarse))
    p.line_search = int(bool(line_search))
    p.history = int(bool(history))
    p.precision = precision
    p.self_contact = int(bool(self_contact))
    p.coarse_solve = int(coarse_solve)
    p.smooth_sweeps = int(smooth_sweeps)
    p.net_level = int(net_level)
    return p


BD_MAX_GRIPS_HELD = 4
BD_GRIP_SLOT_ENTRIES = 512
BD_PICK_MAX_WINDOW_RADIUS = 7


class bd_pick(C.Structure):
    _fields_ = [("node", C.c_int32), ("kind", C.c_int32), ("position", C.c_double * 3), ("window_radius", C.c_int32),
                ("window", C.c_uint32 * ((2 * BD_PICK_MAX_WINDOW_RADIUS + 1) ** 2))]


class GripBuffers:
    """A `bd_grips` struct for a list of grips, each a dict with nodes, weights, offsets (E, 3),
    goal (3,) and stiffness k (N/m per unit weight)."""

    def __init__(self, grips):
        nodes, gid, w, off = [], [], [], []
        goals, ks = [], []
        for g_i, g in enumerate(grips):
            nodes.extend(int(n) for n in g["nodes"])
            gid.extend([g_i] * len(g["nodes"]))
            w.extend(float(x) for x in g["weights"])
            off.append(np.asarray(g["offsets"], dtype=np.float64).reshape(-1, 3))
            goals.append(np.asarray(g["goal"], dtype=np.float64))
            ks.append(float(g["k"]))
        self.node = np.asarray(nodes, dtype=np.int32)
        self.grip = np.asarray(gid, dtype=np.int32)
        self.weight = np.asarray(w, dtype=np.float64)
        self.offset = np.ascontiguousarray(np.concatenate(off) if off else np.zeros((0, 3))).ravel()
        self.goal = np.ascontiguousarray(np.concatenate(goals) if goals else np.zeros(0))
        self.k = np.asarray(ks, dtype=np.float64)
        self.struct = bd_grips(len(grips), len(self.node), _ptr(self.node, C.c_int32), _ptr(self.grip, C.c_int32),
                               _ptr(self.weight, C.c_double), _ptr(self.offset, C.c_double),
                               _ptr(self.goal, C.c_double), _ptr(self.k, C.c_double))


_lib = None


def load_library(path: os.PathLike | None = None):
    """Load libbd.so (built in-tree by the package's build step). Raises if it is missing: there
    is no fallback path for the solve."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = C.CDLL(str(p))
    P = C.POINTER
    vp = C.c_void_p
    lib.bd_last_error.restype = C.c_char_p
    lib.bd_last_error.argtypes = [vp]
    lib.bd_create.argtypes = [P(bd_scene), P(bd_params), P(vp)]
    lib.bd_destroy.argtypes = [vp]
    lib.bd_set_state.argtypes = [vp, P(C.c_double), P(C.c_double)]
    lib.bd_set_segments.argtypes = [vp, P(C.c_double), P(C.c_double), P(C.c_uint32), P(C.c_uint16),
                                    P(C.c_double), P(C.c_double)]
    lib.bd_set_field.argtypes = [vp, P(C.c_uint8)]
    lib.bd_set_grips.argtypes = [vp, P(bd_grips)]
    lib.bd_set_timestep.argtypes = [vp, C.c_double]
    lib.bd_set_history.argtypes = [vp, C.c_int]
    lib.bd_set_newton_tolerance.argtypes = [vp, C.c_double]
    lib.bd_set_pcg.argtypes = [vp, C.c_int, C.c_double, C.c_int]
    lib.bd_substep.argtypes = [vp, C.c_int, C.c_int]
    for name in ("bd_predict", "bd_vcycle", "bd_pcg", "bd_line_search", "bd_finalize", "bd_damage",
                 "bd_slip", "bd_sync"):
        getattr(lib, name).argtypes = [vp]
    lib.bd_assemble.argtypes = [vp, C.c_int]
    lib.bd_smooth.argtypes = [vp, C.c_int, C.c_int]
    lib.bd_residual.argtypes = [vp, C.c_int]
    lib.bd_newton_apply.argtypes = [vp, C.c_int]
    lib.bd_read.argtypes = [vp, C.c_int, C.c_int, vp, C.c_size_t]
    lib.bd_write.argtypes = [vp, C.c_int, C.c_int, vp, C.c_size_t]
    lib.bd_buffer_bytes.argtypes = [vp, C.c_int, C.c_int]
    lib.bd_buffer_bytes.restype = C.c_size_t
    lib.bd_get_stats.argtypes = [vp, P(bd_stats)]
    lib.bd_set_grip_goals.argtypes = [vp, C.c_int, P(C.c_double)]
    lib.bd_profile_enable.argtypes = [vp, C.c_int]
    lib.bd_profile_read.argtypes = [vp, P(bd_kernel_profile), C.c_int, P(C.c_int)]
    for name in ("bd_contact_detect", "bd_contact_terms", "bd_project_start"):
        getattr(lib, name).argtypes = [vp]
    lib.bd_set_floor.argtypes = [vp, C.c_int, C.c_double]
    u32p, i32p, u8p = P(C.c_uint32), P(C.c_int32), P(C.c_uint8)
    f32p, f64p = P(C.c_float), P(C.c_double)
    lib.bd_grip_attach.argtypes = [vp, C.c_int, C.c_int32, i32p, f32p, f32p, C.c_int, C.c_double, C.c_double, C.c_double,
                                   C.c_double, f64p]
    lib.bd_grip_lead.argtypes = [vp]
    lib.bd_grip_terms.argtypes = [vp]
    lib.bd_grip_target.argtypes = [vp, C.c_int, f64p]
    lib.bd_grip_pinch.argtypes = [vp, C.c_int, C.c_double]
    lib.bd_grip_hold.argtypes = [vp, C.c_int, C.c_double]
    lib.bd_grip_release.argtypes = [vp, C.c_int]
    lib.bd_pick_ray.argtypes = [vp, f32p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                f64p, f64p, P(bd_pick)]
    lib.bd_pick_near.argtypes = [vp, f64p, C.c_double, C.c_int, P(bd_pick)]
    lib.bd_autotune.argtypes = [vp, C.c_int, i32p]
    lib.bd_set_smoother.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    lib.bd_set_pieces.argtypes = [vp, u32p, u32p, C.c_int, u32p]
    lib.bd_update_topology.argtypes = [vp, i32p, C.c_int, i32p]
    lib.bd_topology_search.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, u8p, u8p, i32p, C.c_int, u8p, u32p, u32p,
                                       u32p, i32p]
    lib.bd_learn_record.argtypes = [vp, C.c_int]
    lib.bd_learn_train.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_int, f64p]
    lib.bd_set_learned.argtypes = [vp, C.c_int]
    lib.bd_learn_get_weights.argtypes = [vp, f32p, C.c_int]
    lib.bd_learn_set_weights.argtypes = [vp, f32p, C.c_int]
    lib.bd_learn_loss_grad.argtypes = [vp, C.c_int, f64p, f32p, C.c_int]
    lib.bd_learn_read_sample.argtypes = [vp, C.c_int, f32p, f64p, f32p, u8p]
    lib.bd_learn_info.argtypes = [vp, i32p, i32p, i32p, i32p]
    lib.bd_learn_read_fields.argtypes = [vp, f32p]
    lib.bd_apply.argtypes = [vp, C.c_int, vp, vp, vp, vp]
    lib.bd_apply_noalloc.argtypes = [vp, C.c_int, vp, vp, vp, vp]
    lib.bd_set_pcg_mode.argtypes = [vp, C.c_int]
    lib.bd_set_line_colors.argtypes = [vp, C.c_int, C.c_int]
    lib.bd_graph_capture.argtypes = [vp, C.c_int]
    lib.bd_graph_substep.argtypes = [vp]
    lib.bd_graph_release.argtypes = [vp]
    lib.bd_batch_create.argtypes = [P(vp), C.c_int, P(vp)]
    lib.bd_batch_create_concurrent.argtypes = [P(vp), C.c_int, C.c_int, P(vp)]
    for name in ("bd_batch_substep", "bd_batch_sync", "bd_batch_destroy"):
        getattr(lib, name).argtypes = [vp]
    lib.bd_batch_last_error.argtypes = [vp]
    lib.bd_batch_last_error.restype = C.c_char_p
    if path is None:
        _lib = lib
    return lib


# Every entry point include/bd.h declares (checked by the CPU test that loads the library).
EXPORTED = (
    "bd_last_error", "bd_create", "bd_destroy", "bd_set_state", "bd_set_segments", "bd_set_field",
    "bd_set_grips", "bd_set_timestep", "bd_set_history", "bd_set_newton_tolerance", "bd_set_pcg",
    "bd_substep", "bd_predict", "bd_assemble", "bd_vcycle", "bd_smooth", "bd_residual", "bd_pcg",
    "bd_line_search", "bd_newton_apply", "bd_finalize", "bd_damage", "bd_slip", "bd_sync", "bd_read",
    "bd_write", "bd_buffer_bytes", "bd_get_stats", "bd_set_grip_goals", "bd_profile_enable", "bd_profile_read",
    "bd_contact_detect", "bd_contact_terms", "bd_project_start", "bd_set_floor",
    "bd_set_pieces", "bd_update_topology", "bd_topology_search",
    "bd_grip_attach", "bd_grip_lead", "bd_grip_terms", "bd_grip_target", "bd_grip_pinch", "bd_grip_hold",
    "bd_grip_release", "bd_pick_ray", "bd_pick_near", "bd_autotune", "bd_set_smoother",
    "bd_learn_record", "bd_learn_train", "bd_set_learned", "bd_learn_get_weights", "bd_learn_set_weights",
    "bd_learn_loss_grad", "bd_learn_read_sample", "bd_learn_info", "bd_learn_read_fields", "bd_apply",
    "bd_apply_noalloc", "bd_set_pcg_mode", "bd_set_line_colors", "bd_graph_capture", "bd_graph_substep", "bd_graph_release", "bd_batch_create",
    "bd_batch_create_concurrent", "bd_batch_substep", "bd_batch_sync", "bd_batch_destroy", "bd_batch_last_error",
)

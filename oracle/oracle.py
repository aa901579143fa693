"""Python handle on the FP64 CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's cpu_baseline leg import this module, as the checker or
the timed CPU baseline; the product path never does. The oracle restates the reference's f64
solver (src/simulation/reference/solver.ts, newtonMultigrid.ts) and its laws and yarn field
(axial.ts, laws.ts, yarnField.ts); each C function cites the reference lines it follows.

Parity pinning: every oracle function is checked in tests/test_oracle_*.py against the
reference's own unit-test invariants and fixtures (axial/laws/newtonMultigrid/referenceSolver/
yarnField .test.ts), closed forms, brute force on tiny inputs and the sceneHash pin.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
sys.path.insert(0, str(ROOT / "break-that-denim_amd"))

from bd.capi import SceneBuffers, GripBuffers, bd_params, bd_scene, bd_grips, bd_yarn, make_params  # noqa: E402

# BD_ORACLE_LIB names another build of the same sources (the sanitizer test's instrumented one)
LIB = Path(os.environ["BD_ORACLE_LIB"]) if os.environ.get("BD_ORACLE_LIB") else HERE / "liboracle.so"
_lib = None

STENCIL = 25
CENTER = 12


class OcScene(C.Structure):
    """oc_scene of contact_oracle.c: node-major FP32 arrays of one scene state."""
    _fields_ = [("N", C.c_int), ("nl", C.c_int), ("nx", C.c_int * 8), ("ny", C.c_int * 8), ("off", C.c_int * 8),
                ("x", C.POINTER(C.c_float)), ("dx", C.POINTER(C.c_float)), ("flags", C.POINTER(C.c_uint16)),
                ("pinned", C.POINTER(C.c_uint8)), ("mass", C.POINTER(C.c_float))]


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        vp = C.c_void_p
        P = C.POINTER
        dp = P(C.c_double)
        L.or_set_threads.argtypes = [C.c_int]
        L.or_set_threads.restype = C.c_int
        L.or_create.restype = vp
        L.or_create.argtypes = [P(bd_scene), P(bd_params)]
        L.or_destroy.argtypes = [vp]
        L.or_set_grips.argtypes = [vp, P(bd_grips)]
        L.or_set_params.argtypes = [vp, P(bd_params)]
        for name in ("or_predict", "or_assemble", "or_finalize", "or_update_damage", "or_update_slip", "or_update_field"):
            getattr(L, name).argtypes = [vp]
        for name in ("or_build_hierarchy", "or_couple_spans", "or_pull_out", "or_level_count", "or_substep",
                     "or_pcg_iterations"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = C.c_int
        L.or_set_ext.argtypes = [vp, dp]
        L.or_set_cmeta.argtypes = [vp, dp]
        L.or_set_curve.argtypes = [vp, dp]
        L.or_set_floor.argtypes = [vp, C.c_int, C.c_double]
        L.or_project_start.argtypes = [vp]
        # FP32 self-contact restatement (contact_oracle.c)
        fp = P(C.c_float)
        up = P(C.c_uint32)
        L.oc_detect.argtypes = [P(OcScene), up, fp, up, fp]
        L.oc_detect_brute.argtypes = [P(OcScene), up, fp]
        L.oc_terms.argtypes = [P(OcScene), up, fp, C.c_float, C.c_float, P(C.c_int32), fp]
        L.oc_closest_tri.argtypes = [fp, fp, fp, fp, fp]
        L.oc_closest_tri.restype = C.c_float
        L.oc_face_hit.argtypes = [fp, fp, fp, fp, fp, fp, fp]
        L.oc_face_hit.restype = C.c_float
        L.oc_seg_tri.argtypes = [fp, fp, fp, fp, fp]
        L.oc_seg_tri.restype = C.c_float
        L.oc_seg_dist.argtypes = [fp, fp, fp]
        L.oc_seg_dist.restype = C.c_float
        L.oc_hash_cell.argtypes = [C.c_int, C.c_int, C.c_int]
        L.oc_hash_cell.restype = C.c_uint32
        L.oc_face_valid.argtypes = [P(OcScene), C.c_uint32]
        L.oc_face_valid.restype = C.c_int
        L.oc_feasible.argtypes = [fp, fp]
        L.oc_floor_bound_z.argtypes = [C.c_float] * 7
        L.oc_floor_bound_z.restype = C.c_float
        L.oc_project_start.argtypes = [P(OcScene), fp, fp, C.c_int, C.c_float, fp, C.c_float]
        # FP32 cross-layer contact restatement (contact_oracle.c)
        L.oc_ply_detect.argtypes = [P(OcScene), fp, up, up, up, fp]
        L.oc_ply_terms.argtypes = [P(OcScene), fp, up, C.c_float, fp]
        L.oc_ply_hash_cell.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint32]
        L.oc_ply_hash_cell.restype = C.c_uint32
        L.or_rhs.argtypes = [vp, dp]
        L.or_node_gradient.argtypes = [vp, C.c_int, dp]
        L.or_potential.argtypes = [vp]
        L.or_potential.restype = C.c_double
        L.or_energy_change.argtypes = [vp, dp, dp, dp]
        L.or_choose_step.argtypes = [dp, dp]
        L.or_choose_step.restype = C.c_double
        L.or_load_fine_operator.argtypes = [vp, dp, P(C.c_uint8)]
        L.or_load_fine_operator.restype = C.c_int
        L.or_level_nodes.argtypes = [vp, C.c_int]
        L.or_level_nodes.restype = C.c_int
        L.or_level_op.argtypes = [vp, C.c_int]
        L.or_level_op.restype = dp
        L.or_level_fixed.argtypes = [vp, C.c_int]
        L.or_level_fixed.restype = P(C.c_uint8)
        L.or_level_pidx.argtypes = [vp, C.c_int]
        L.or_level_pidx.restype = P(C.c_int)
        L.or_level_pw.argtypes = [vp, C.c_int]
        L.or_level_pw.restype = dp
        L.or_apply.argtypes = [vp, C.c_int, dp, dp]
        L.or_fp16_round.argtypes = [C.c_double]
        L.or_fp16_round.restype = C.c_double
        L.or_smooth.argtypes = [vp, C.c_int, dp, dp, C.c_int]
        L.or_set_line_colors.argtypes = [vp, C.c_int, C.c_int]
        L.or_vcycle.argtypes = [vp, dp, dp]
        L.or_pcg.argtypes = [vp, dp, dp, C.c_int, C.c_double, C.c_int]
        L.or_pcg_relres.argtypes = [vp]
        L.or_pcg_relres.restype = C.c_double
        L.or_newton_step.argtypes = [vp, C.c_int]
        L.or_newton_step.restype = C.c_double
        L.or_yarn_material_length.argtypes = [vp]
        L.or_yarn_material_length.restype = C.c_double
        L.or_ptr.argtypes = [vp, C.c_int]
        L.or_ptr.restype = vp
        L.or_counter.argtypes = [vp, C.c_int]
        L.or_counter.restype = C.c_int
        L.or_last_alpha.argtypes = [vp]
        L.or_last_alpha.restype = C.c_double
        L.or_solve_penta.argtypes = [C.c_int, dp, dp, dp, dp, dp, C.c_double]
        yp = P(bd_yarn)
        for name, args in (("or_axial_tension", [yp, C.c_double]), ("or_axial_stiffness", [yp, C.c_double]),
                           ("or_axial_energy_density", [yp, C.c_double]), ("or_strain_at_tension", [yp, C.c_double]),
                           ("or_yarn_damage", [yp] + [C.c_double] * 5),
                           ("or_bend_psi_delta", [C.c_double] * 4), ("or_shear_jam", [C.c_double]),
                           ("or_cell_area", [C.c_double] * 4),
                           ("or_bend_scale", [C.c_double, C.c_double, C.c_uint32, C.c_uint32, C.c_int]),
                           ("or_recoil_length", [yp] + [C.c_double] * 9)):
            getattr(L, name).argtypes = args
            getattr(L, name).restype = C.c_double
        L.or_interpolated_strain.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_interpolated_strain.restype = C.c_double
        L.or_share_factor.argtypes = [vp, C.c_int, P(C.c_uint8), C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        L.or_share_factor.restype = C.c_double
        L.or_span_crossing.argtypes = [vp] + [C.c_int] * 6
        L.or_span_crossing.restype = C.c_int
        L.or_share_reach.argtypes = [vp, C.c_int, C.c_int]
        L.or_share_reach.restype = C.c_int
        L.or_share_cap.argtypes = [C.c_int] + [C.c_double] * 5
        L.or_share_cap.restype = C.c_double
        L.or_bend_psi.argtypes = [C.c_double, C.c_double, dp]
        L.or_bend_geometry.argtypes = [C.c_double, C.c_double, dp]
        L.or_shear_law.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def yarn(y) -> bd_yarn:
    from bd.capi import _yarn_struct
    return _yarn_struct(y)


# ---- law helpers (scalar)
def axial_tension(y, eps): return lib().or_axial_tension(C.byref(yarn(y)), eps)
def axial_stiffness(y, eps): return lib().or_axial_stiffness(C.byref(yarn(y)), eps)
def axial_energy_density(y, eps): return lib().or_axial_energy_density(C.byref(yarn(y)), eps)
def strain_at_tension(y, t): return lib().or_strain_at_tension(C.byref(yarn(y)), t)
def yarn_damage(y, B, area, peak, soft, eps): return lib().or_yarn_damage(C.byref(yarn(y)), B, area, peak, soft, eps)
def bend_psi_delta(q0, q1, dq, a): return lib().or_bend_psi_delta(q0, q1, dq, a)
def shear_jam(area): return lib().or_shear_jam(area)
def cell_area(cu, tu, cv, rv): return lib().or_cell_area(cu, tu, cv, rv)
def bend_scale(ia, ic, ma, mc, B): return lib().or_bend_scale(ia, ic, ma, mc, B)
def recoil_length(y, B, area, peak, eps, mu, n0, wrap, tc, pitch):
    return lib().or_recoil_length(C.byref(yarn(y)), B, area, peak, eps, mu, n0, wrap, tc, pitch)


def bend_psi(q, a):
    o = np.zeros(3)
    lib().or_bend_psi(q, a, _dp(o))
    return o


def bend_geometry(La, Lc):
    o = np.zeros(4)
    lib().or_bend_geometry(La, Lc, _dp(o))
    return {"wa": o[0], "wc": o[1], "kPerB": o[2], "fold": o[3]}


def shear_law(ks, Ld, l, jam=0.16):
    o = np.zeros(3)
    lib().or_shear_law(ks, Ld, l, jam, _dp(o))
    return o


def choose_step(e, m):
    e = np.ascontiguousarray(e, dtype=np.float64)
    m = np.ascontiguousarray(m, dtype=np.float64)
    return lib().or_choose_step(_dp(e), _dp(m))


def solve_penta(D, E, F, r, shift=0.0):
    n = len(r) // 3
    D, E, F, r = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (D, E, F, r))
    x = np.zeros(3 * n)
    lib().or_solve_penta(n, _dp(D), _dp(E), _dp(F), _dp(r), _dp(x), shift)
    return x


class Oracle:
    """CPU FP64 twin of one solver handle on a compiled scene."""

    _PTR = {"x": (0, np.float64, 3), "v": (1, np.float64, 3), "dx": (2, np.float64, 3), "pred": (3, np.float64, 3),
            "rest": (4, np.float64, 2), "intact": (5, np.float64, 2), "cut_mask": (6, np.uint32, 2),
            "seg_flags": (7, np.uint16, 2), "fail_site": (8, np.float64, 2), "tails": (9, np.float64, 4),
            "last_update": (11, np.float64, 1), "strain": (12, np.float64, 2), "pinned": (15, np.uint8, 1)}

    def __init__(self, scene, params=None, grips=None):
        self.scene = scene
        self.N = scene.node_count
        self.params = params if params is not None else make_params()
        self._sb = SceneBuffers(scene)
        self.h = lib().or_create(C.byref(self._sb.scene), C.byref(self.params))
        self._grips = None
        if grips:
            self.set_grips(grips)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def view(self, name):
        which, dt, per = self._PTR[name]
        p = lib().or_ptr(self.h, which)
        n = per * self.N
        buf = (C.c_byte * (n * np.dtype(dt).itemsize)).from_address(p)
        return np.frombuffer(buf, dtype=dt, count=n)

    def __getattr__(self, name):
        if name in Oracle._PTR:
            return self.view(name)
        raise AttributeError(name)

    @property
    def damage(self):
        p = lib().or_ptr(self.h, 10)
        n = 2 * self.scene.crossing_count
        return np.frombuffer((C.c_uint8 * n).from_address(p), dtype=np.uint8, count=n)

    def set_params(self, params):
        self.params = params
        lib().or_set_params(self.h, C.byref(params))

    def set_grips(self, grips):
        self._grips = GripBuffers(grips) if grips else None
        lib().or_set_grips(self.h, C.byref(self._grips.struct) if self._grips else None)

    def predict(self): lib().or_predict(self.h)
    def assemble(self): lib().or_assemble(self.h)
    def build_hierarchy(self): return lib().or_build_hierarchy(self.h)
    def finalize(self): lib().or_finalize(self.h)
    def update_damage(self): lib().or_update_damage(self.h)
    def update_slip(self): lib().or_update_slip(self.h)
    def update_field(self): lib().or_update_field(self.h)
    def pull_out(self): return lib().or_pull_out(self.h)
    def couple_spans(self): return lib().or_couple_spans(self.h)
    def substep(self): return lib().or_substep(self.h)
    def potential(self): return lib().or_potential(self.h)
    def newton_step(self, first=True): return lib().or_newton_step(self.h, int(first))
    def yarn_material_length(self): return lib().or_yarn_material_length(self.h)

    @property
    def broken_count(self): return lib().or_counter(self.h, 0)

    @property
    def pulled_out_count(self): return lib().or_counter(self.h, 1)

    @property
    def rejections(self): return lib().or_counter(self.h, 2)

    @property
    def last_alpha(self): return lib().or_last_alpha(self.h)

    @property
    def pcg_iterations(self): return lib().or_pcg_iterations(self.h)

    @property
    def pcg_relres(self): return lib().or_pcg_relres(self.h)

    # ---- external terms, contact bounds, floor (FP64 restatement of their use)
    def set_ext(self, ext):
        """(N, 9) gradient + Hessian [xx xy xz yy yz zz] per node, or None."""
        a = None if ext is None else np.ascontiguousarray(ext, dtype=np.float64).ravel()
        lib().or_set_ext(self.h, _dp(a) if a is not None else None)

    def set_curve(self, curve):
        """Detached pieces' curved line-search steps, (N, 7): rotation w xyz, flag, arm r xyz (or None)."""
        if curve is None:
            lib().or_set_curve(self.h, None)
            return
        c = np.ascontiguousarray(curve, dtype=np.float64).reshape(-1)
        lib().or_set_curve(self.h, c.ctypes.data_as(C.POINTER(C.c_double)))

    def set_cmeta(self, cmeta):
        """(N, 2, 4) own and reciprocal bounds, or None."""
        a = None if cmeta is None else np.ascontiguousarray(cmeta, dtype=np.float64).ravel()
        lib().or_set_cmeta(self.h, _dp(a) if a is not None else None)

    def set_floor(self, z):
        lib().or_set_floor(self.h, 0 if z is None else 1, 0.0 if z is None else float(z))

    def project_start(self):
        lib().or_project_start(self.h)

    def rhs(self):
        b = np.zeros(3 * self.N)
        lib().or_rhs(self.h, _dp(b))
        return b

    def node_gradient(self, n):
        g = np.zeros(3)
        lib().or_node_gradient(self.h, int(n), _dp(g))
        return g

    def energy_change(self, d):
        d = np.ascontiguousarray(d, dtype=np.float64).ravel()
        e = np.zeros(4)
        m = np.zeros(4)
        lib().or_energy_change(self.h, _dp(d), _dp(e), _dp(m))
        return e, m

    # ---- multigrid levels
    def level_count(self): return lib().or_level_count(self.h)

    def level_nodes(self, k): return lib().or_level_nodes(self.h, k)

    def stencil(self, k=0):
        n = self.level_nodes(k)
        p = lib().or_level_op(self.h, k)
        return np.ctypeslib.as_array(p, shape=(n * STENCIL * 9,)).reshape(n, STENCIL, 3, 3).copy()

    def fixed(self, k=0):
        n = self.level_nodes(k)
        p = lib().or_level_fixed(self.h, k)
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def prolongation(self, k=0):
        n = self.level_nodes(k)
        idx = np.ctypeslib.as_array(lib().or_level_pidx(self.h, k), shape=(4 * n,)).reshape(n, 4).copy()
        w = np.ctypeslib.as_array(lib().or_level_pw(self.h, k), shape=(4 * n,)).reshape(n, 4).copy()
        return idx, w

    def interpolated_strain(self, f, k, m, layer=0):
        return lib().or_interpolated_strain(self.h, layer, f, k, m)

    def share_factor(self, sep, f, k, m, reach, cap=float("inf"), layer=0):
        sep = np.ascontiguousarray(sep, dtype=np.uint8)
        return lib().or_share_factor(self.h, layer, sep.ctypes.data_as(C.POINTER(C.c_uint8)), f, k, m, reach, cap)

    def span_crossing(self, i, j, f, q, p, layer=0):
        return lib().or_span_crossing(self.h, layer, i, j, f, q, p)

    def share_reach(self, f, layer=0):
        return lib().or_share_reach(self.h, layer, f)

    def load_fine_operator(self, A, fixed):
        A = np.ascontiguousarray(A, dtype=np.float64).ravel()
        fixed = np.ascontiguousarray(fixed, dtype=np.uint8)
        rc = lib().or_load_fine_operator(self.h, _dp(A), fixed.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc != 0:
            raise RuntimeError("galerkin stencil exceeds radius 2")

    def apply(self, k, x):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        y = np.zeros_like(x)
        lib().or_apply(self.h, k, _dp(x), _dp(y))
        return y

    def set_line_colors(self, k, colors):
        """Line colours per family of level k's smoother (newtonMultigrid.ts lineColors; 1 is an extension)."""
        if lib().or_set_line_colors(self.h, int(k), int(colors)) != 0:
            raise ValueError("or_set_line_colors: level or colours out of range")

    def smooth(self, k, x, b, reverse=False):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel().copy()
        b = np.ascontiguousarray(b, dtype=np.float64).ravel()
        lib().or_smooth(self.h, k, _dp(x), _dp(b), int(reverse))
        return x

    def vcycle(self, b, x=None):
        b = np.ascontiguousarray(b, dtype=np.float64).ravel()
        x = np.zeros_like(b) if x is None else np.ascontiguousarray(x, dtype=np.float64).ravel().copy()
        lib().or_vcycle(self.h, _dp(x), _dp(b))
        return x

    def pcg(self, b, iterations, rel_tol=0.0, max_iter=200, x=None):
        b = np.ascontiguousarray(b, dtype=np.float64).ravel()
        x = np.zeros_like(b) if x is None else np.ascontiguousarray(x, dtype=np.float64).ravel().copy()
        lib().or_pcg(self.h, _dp(b), _dp(x), iterations, rel_tol, max_iter)
        return x


class ContactTwin:
    """The FP32 restatement of the reference's self-contact (contact_oracle.c) on one scene state:
    x, dx (N, 3) float32 and the span flags; K = 4 candidates per node."""

    K = 4
    NONE = 0xFFFFFFFF
    TABLE = 1 << 17

    def __init__(self, scene, x, dx, flags=None, pinned=None, mass=None):
        self.N = scene.node_count
        L = scene.layers
        self.keep = {
            "x": np.ascontiguousarray(x, dtype=np.float32).reshape(-1),
            "dx": np.ascontiguousarray(dx, dtype=np.float32).reshape(-1),
            "flags": np.ascontiguousarray(scene.seg_flags if flags is None else flags, dtype=np.uint16),
            "pinned": np.ascontiguousarray(scene.pinned if pinned is None else pinned, dtype=np.uint8),
            "mass": np.ascontiguousarray(scene.mass if mass is None else mass, dtype=np.float32),
        }
        k = self.keep
        S = OcScene()
        S.N = self.N
        S.nl = len(L)
        for i, lay in enumerate(L):
            S.nx[i], S.ny[i], S.off[i] = lay.nx, lay.ny, lay.node_offset
        S.x = k["x"].ctypes.data_as(C.POINTER(C.c_float))
        S.dx = k["dx"].ctypes.data_as(C.POINTER(C.c_float))
        S.flags = k["flags"].ctypes.data_as(C.POINTER(C.c_uint16))
        S.pinned = k["pinned"].ctypes.data_as(C.POINTER(C.c_uint8))
        S.mass = k["mass"].ctypes.data_as(C.POINTER(C.c_float))
        self.S = S
        self.half = np.array([0.5 * lay.material.get("thickness", 0.0) for lay in L] + [0.0] * (8 - len(L)), dtype=np.float32)

    def detect(self):
        """faceCount .. resolveBounds: returns cells (2T,), items (2N, 8), pairs (N, K+2), cmeta (N, 2, 4)."""
        cells = np.zeros(2 * self.TABLE, dtype=np.uint32)
        items = np.zeros((2 * self.N, 8), dtype=np.float32)
        pairs = np.zeros((self.N, self.K + 2), dtype=np.uint32)
        cmeta = np.zeros((self.N, 2, 4), dtype=np.float32)
        lib().oc_detect(C.byref(self.S), _up(cells), _fp(items), _up(pairs), _fp(cmeta))
        return cells, items, pairs, cmeta

    def detect_brute(self):
        pairs = np.zeros((self.N, self.K + 2), dtype=np.uint32)
        cmeta = np.zeros((self.N, 2, 4), dtype=np.float32)
        lib().oc_detect_brute(C.byref(self.S), _up(pairs), _fp(cmeta))
        return pairs, cmeta

    def terms(self, pairs, cmeta, h, gnorm):
        """terms + gather: returns (acc (N, 9) int32, ext (N, 9) float32); cmeta own bounds tightened in place."""
        acc = np.zeros((self.N, 9), dtype=np.int32)
        ext = np.zeros((self.N, 9), dtype=np.float32)
        pr = np.ascontiguousarray(pairs, dtype=np.uint32)
        assert cmeta.dtype == np.float32 and cmeta.flags.c_contiguous
        lib().oc_terms(C.byref(self.S), _up(pr), _fp(cmeta), C.c_float(h), C.c_float(gnorm),
                       acc.ctypes.data_as(C.POINTER(C.c_int32)), _fp(ext))
        return acc, ext

    def project_start(self, cmeta, floor_z=None, gnorm=9.81):
        """projectStart on a copy of dx: returns the feasible predicted step (N, 3)."""
        dx = self.keep["dx"].copy()
        lib().oc_project_start(C.byref(self.S), _fp(np.ascontiguousarray(cmeta, dtype=np.float32)), _fp(dx),
                               0 if floor_z is None else 1, C.c_float(0.0 if floor_z is None else floor_z),
                               _fp(self.half), C.c_float(gnorm))
        return dx.reshape(-1, 3)

    def face_valid(self, f):
        return bool(lib().oc_face_valid(C.byref(self.S), int(f)))


class PlyTwin(ContactTwin):
    """The FP32 restatement of the reference's cross-layer contact (contact.wgsl, in contact_oracle.c)
    on one scene state of 2-3 layers: K nearest foreign nodes per node, bounds and terms."""

    TABLE = 1 << 20

    def __init__(self, scene, x, dx, **kw):
        super().__init__(scene, x, dx, **kw)
        t = [lay.material.get("thickness", 0.0) for lay in scene.layers] + [0.0, 0.0]
        # gpu/contact.ts:52-59: half the summed thicknesses, in f64, stored as f32
        self.dist3 = np.array([0.5 * (t[0] + t[1]), 0.5 * (t[0] + t[2]), 0.5 * (t[1] + t[2])], dtype=np.float32)

    def detect(self):
        """hashCount .. detect: returns cells (2T,), items (N,), pairs (N, K), cmeta (N, 2, 4)."""
        cells = np.zeros(2 * self.TABLE, dtype=np.uint32)
        items = np.zeros(max(self.N, 1), dtype=np.uint32)
        pairs = np.zeros((self.N, self.K), dtype=np.uint32)
        cmeta = np.zeros((self.N, 2, 4), dtype=np.float32)
        lib().oc_ply_detect(C.byref(self.S), _fp(self.dist3), _up(cells), _up(items), _up(pairs), _fp(cmeta))
        return cells, items[:self.N], pairs, cmeta

    def terms(self, pairs, h):
        """terms: returns ext (N, 9) float32 (gradient xyz, Hessian xx xy xz yy yz zz)."""
        ext = np.zeros((self.N, 9), dtype=np.float32)
        pr = np.ascontiguousarray(pairs, dtype=np.uint32)
        lib().oc_ply_terms(C.byref(self.S), _fp(self.dist3), _up(pr), C.c_float(h), _fp(ext))
        return ext


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _f3(v):
    return np.ascontiguousarray(v, dtype=np.float32)


def closest_tri(p, a, b, c):
    w = np.zeros(3, dtype=np.float32)
    d2 = lib().oc_closest_tri(_fp(_f3(p)), _fp(_f3(a)), _fp(_f3(b)), _fp(_f3(c)), _fp(w))
    return w, d2


def face_hit(p, p00, p10, p01, p11):
    w = np.zeros(4, dtype=np.float32)
    q = np.zeros(3, dtype=np.float32)
    d = lib().oc_face_hit(_fp(_f3(p)), _fp(_f3(p00)), _fp(_f3(p10)), _fp(_f3(p01)), _fp(_f3(p11)), _fp(w), _fp(q))
    return w, q, d


def seg_tri(o, d, a, b, c):
    return lib().oc_seg_tri(_fp(_f3(o)), _fp(_f3(d)), _fp(_f3(a)), _fp(_f3(b)), _fp(_f3(c)))


def seg_dist(q, a, b):
    return lib().oc_seg_dist(_fp(_f3(q)), _fp(_f3(a)), _fp(_f3(b)))


def hash_cell(cx, cy, cz):
    return lib().oc_hash_cell(cx, cy, cz)


def feasible(cmeta_n, d):
    dd = _f3(d).copy()
    lib().oc_feasible(_fp(np.ascontiguousarray(cmeta_n, dtype=np.float32).ravel()), _fp(dd))
    return dd


def floor_bound_z(z0, linz, dz, half, mass, gnorm, fz):
    return lib().oc_floor_bound_z(z0, linz, dz, half, mass, gnorm, fz)


def fp16_round(a):
    """The oracle's FP16 rounding of one value (round to nearest even, saturated at 65504)."""
    return lib().or_fp16_round(float(a))


def set_threads(n=0):
    """Threads of the oracle's row and line loops (0 keeps the current count); returns the count in
    use. Results do not depend on it."""
    return lib().or_set_threads(int(n))

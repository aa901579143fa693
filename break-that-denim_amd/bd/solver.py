"""Host-side mirror of the reference's GpuSolver (src/simulation/gpu/solver.ts:194-838) over libbd.so.

Method names follow the reference's (encodeSubstep -> encode_substep, setState -> set_state, ...);
the stage encoders enqueue on the solver's HIP stream and return, readbacks synchronise. Failing
calls raise RuntimeError with the library's message, as the reference throws JS Errors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import (BD_BUF_CELLS, BD_BUF_CMETA, BD_BUF_EXT, BD_BUF_ITEMS, BD_BUF_PAIRS, BD_CONTACT_K,
                   BD_BUF_CUTMASK, BD_BUF_DELTA, BD_BUF_DX, BD_BUF_ENERGY, BD_BUF_FAILSITE, BD_BUF_FIELD,
                   BD_BUF_FIXED, BD_BUF_INTACT, BD_BUF_LAST_UPDATE, BD_BUF_LEVEL_B, BD_BUF_LEVEL_R, BD_BUF_LEVEL_X,
                   BD_BUF_PRED, BD_BUF_PROLONG, BD_BUF_REST, BD_BUF_RHS, BD_BUF_SEGFLAGS, BD_BUF_STENCIL,
                   BD_BUF_TAILS, BD_BUF_V, BD_BUF_X, BD_BUF_STATE, BD_BUF_RIGIDK, BD_BUF_PIECES, BD_BUF_CURVE,
                   GripBuffers, SceneBuffers, bd_stats, make_params)

_DTYPES = {
    BD_BUF_FIXED: np.uint8, BD_BUF_CUTMASK: np.uint32, BD_BUF_SEGFLAGS: np.uint16, BD_BUF_FIELD: np.uint8,
    BD_BUF_EXT: np.float32, BD_BUF_CMETA: np.float32, BD_BUF_PAIRS: np.uint32, BD_BUF_CELLS: np.uint32,
    BD_BUF_ITEMS: np.float32, BD_BUF_STATE: np.uint8, BD_BUF_RIGIDK: np.float32, BD_BUF_CURVE: np.float32,
}


class GpuSolver:
    """One implicit-solve handle on the GPU (gpu/solver.ts:275 constructor)."""

    def __init__(self, scene, params=None, grips=None):
        self.lib = capi.load_library()
        self.scene = scene
        self.N = scene.node_count
        self.params = params if params is not None else make_params()
        self._sb = SceneBuffers(scene)
        h = C.c_void_p()
        rc = self.lib.bd_create(C.byref(self._sb.scene), C.byref(self.params), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"bd_create failed ({rc}): {self.lib.bd_last_error(None).decode()}")
        self.h = h
        # a scene of several plies has cross-layer contact; a one-layer scene self-contact when asked
        # for (gpu/solver.ts:322, 364)
        self.contact_mode = "plies" if len(scene.layers) > 1 else ("self" if self.params.self_contact else None)
        self._grips = None
        if grips:
            self.set_grips(grips)

    def close(self):
        if getattr(self, "h", None):
            self.lib.bd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self.lib.bd_last_error(self.h).decode()}")

    # ---- state upload (gpu/solver.ts:801-837)
    def set_state(self, position, velocity=None):
        x = np.ascontiguousarray(position, dtype=np.float64).ravel()
        v = None if velocity is None else np.ascontiguousarray(velocity, dtype=np.float64).ravel()
        self._check(self.lib.bd_set_state(self.h, x.ctypes.data_as(C.POINTER(C.c_double)),
                                          None if v is None else v.ctypes.data_as(C.POINTER(C.c_double))), "bd_set_state")

    def set_segments(self, rest=None, intact=None, cut_mask=None, seg_flags=None, fail_site=None, tails=None):
        def p(a, dt, ct):
            if a is None:
                return None, None
            arr = np.ascontiguousarray(a, dtype=dt)
            return arr, arr.ctypes.data_as(C.POINTER(ct))
        keep = [p(rest, np.float64, C.c_double), p(intact, np.float64, C.c_double), p(cut_mask, np.uint32, C.c_uint32),
                p(seg_flags, np.uint16, C.c_uint16), p(fail_site, np.float64, C.c_double), p(tails, np.float64, C.c_double)]
        self._check(self.lib.bd_set_segments(self.h, *[k[1] for k in keep]), "bd_set_segments")

    def set_field(self, damage):
        d = np.ascontiguousarray(damage, dtype=np.uint8)
        self._check(self.lib.bd_set_field(self.h, d.ctypes.data_as(C.POINTER(C.c_uint8))), "bd_set_field")

    def set_grips(self, grips):
        self._grips = GripBuffers(grips) if grips else None
        self._check(self.lib.bd_set_grips(self.h, C.byref(self._grips.struct) if self._grips else None), "bd_set_grips")

    def set_grip_goals(self, goals):
        """Move the current grips' goals (one xyz per grip) without re-uploading the patches."""
        g = np.ascontiguousarray(goals, dtype=np.float64).reshape(-1, 3)
        self._check(self.lib.bd_set_grip_goals(self.h, len(g), g.ctypes.data_as(C.POINTER(C.c_double))),
                    "bd_set_grip_goals")

    def profile_enable(self, on=True):
        self._check(self.lib.bd_profile_enable(self.h, int(bool(on))), "bd_profile_enable")

    def profile_read(self):
        """{kernel class: (launches, total ms, algorithmic bytes)} since the last read."""
        arr = (capi.bd_kernel_profile * 32)()
        n = C.c_int()
        self._check(self.lib.bd_profile_read(self.h, arr, 32, C.byref(n)), "bd_profile_read")
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms, arr[i].bytes) for i in range(n.value)}

    def set_timestep(self, h):
        self._check(self.lib.bd_set_timestep(self.h, h), "bd_set_timestep")

    def set_history_enabled(self, on):
        self._check(self.lib.bd_set_history(self.h, int(bool(on))), "bd_set_history")

    def set_newton_tolerance(self, tol):
        self._check(self.lib.bd_set_newton_tolerance(self.h, tol), "bd_set_newton_tolerance")

    def set_pcg(self, iterations=2, rel_tol=0.0, max_iterations=200):
        self._check(self.lib.bd_set_pcg(self.h, iterations, rel_tol, max_iterations), "bd_set_pcg")

    # ---- schedule (gpu/solver.ts:574-760)
    def encode_substep(self, min_iterations=0, max_iterations=0):
        self._check(self.lib.bd_substep(self.h, min_iterations, max_iterations), "bd_substep")

    substep = encode_substep

    def set_line_colors(self, level, colors):
        """Line colours per family of one level's smoother: 3 (the reference), 2 (lines two apart
        together, their couplings majorised; split kernels only) or 1 (every line of a family at once
        against the others' lagged iterate), bd_set_line_colors."""
        self._check(self.lib.bd_set_line_colors(self.h, int(level), int(colors)), "bd_set_line_colors")

    def set_pcg_mode(self, mode):
        """BD_PCG_GRID (0): PCG kernels over the chip; BD_PCG_WORKGROUP (1): the whole solve in one
        workgroup (small systems, batches of patches)."""
        self._check(self.lib.bd_set_pcg_mode(self.h, int(mode)), "bd_set_pcg_mode")

    # ---- graph-captured substep (encodeSubstep recorded once, replayed per frame)
    def graph_capture(self, pcg_iterations):
        """Record one substep into a HIP graph; each Newton solve gets `pcg_iterations` PCG iterations,
        those past convergence gated off on the device."""
        self._check(self.lib.bd_graph_capture(self.h, int(pcg_iterations)), "bd_graph_capture")

    def graph_substep(self):
        self._check(self.lib.bd_graph_substep(self.h), "bd_graph_substep")

    def graph_release(self):
        self._check(self.lib.bd_graph_release(self.h), "bd_graph_release")

    def encode_predict(self):
        self._check(self.lib.bd_predict(self.h), "bd_predict")

    def encode_assemble(self, coarse=True):
        self._check(self.lib.bd_assemble(self.h, int(bool(coarse))), "bd_assemble")

    def encode_vcycle(self):
        self._check(self.lib.bd_vcycle(self.h), "bd_vcycle")

    def encode_smooth(self, level, reverse=False):
        self._check(self.lib.bd_smooth(self.h, level, int(bool(reverse))), "bd_smooth")

    def encode_residual(self, level):
        self._check(self.lib.bd_residual(self.h, level), "bd_residual")

    def encode_pcg(self):
        self._check(self.lib.bd_pcg(self.h), "bd_pcg")

    def encode_line_search(self):
        self._check(self.lib.bd_line_search(self.h), "bd_line_search")

    def encode_newton_apply(self, iteration=0):
        self._check(self.lib.bd_newton_apply(self.h, iteration), "bd_newton_apply")

    def encode_finalize(self):
        self._check(self.lib.bd_finalize(self.h), "bd_finalize")

    def encode_damage(self):
        self._check(self.lib.bd_damage(self.h), "bd_damage")

    def encode_slip(self):
        self._check(self.lib.bd_slip(self.h), "bd_slip")

    def write_velocity(self, v):
        """Overwrite the velocities (3N), e.g. the kinetic damping of the drape settle."""
        self.write(BD_BUF_V, np.asarray(v, dtype=np.float64).ravel())

    def encode_contact_detect(self):
        """SelfContact.encodeDetect (gpu/selfcontact.ts:139-150), or Contact.encodeDetect
        (gpu/contact.ts:129-137) for a scene of several plies."""
        self._check(self.lib.bd_contact_detect(self.h), "bd_contact_detect")

    def encode_contact_terms(self):
        """SelfContact.encodeTerms (gpu/selfcontact.ts:153-157), or Contact.encodeTerms
        (gpu/contact.ts:139-142) for a scene of several plies."""
        self._check(self.lib.bd_contact_terms(self.h), "bd_contact_terms")

    def encode_project_start(self):
        """projectStart (step.wgsl:85-95)."""
        self._check(self.lib.bd_project_start(self.h), "bd_project_start")

    def set_floor(self, z):
        """setFloor(z | null) (gpu/solver.ts:796-801)."""
        self._check(self.lib.bd_set_floor(self.h, 0 if z is None else 1, 0.0 if z is None else float(z)), "bd_set_floor")

    # ---- detached pieces (topology.ts, rigid_*.wgsl)
    def set_pieces(self, piece_of, start, nodes):
        """GpuSolver.setPieces (gpu/solver.ts:664-676)."""
        of = np.ascontiguousarray(piece_of, dtype=np.uint32)
        st = np.ascontiguousarray(start, dtype=np.uint32)
        nd = np.ascontiguousarray(nodes if len(nodes) else [0], dtype=np.uint32)
        u32 = C.POINTER(C.c_uint32)
        self._check(self.lib.bd_set_pieces(self.h, of.ctypes.data_as(u32), st.ctypes.data_as(u32), len(st) - 1,
                                           nd.ctypes.data_as(u32)), "bd_set_pieces")

    def update_topology(self, anchors=()):
        """Topology.request (topology.ts:98-160), synchronously: returns the number of nodes freed."""
        a = np.ascontiguousarray(anchors if len(anchors) else [0], dtype=np.int32)
        out = C.c_int32()
        self._check(self.lib.bd_update_topology(self.h, a.ctypes.data_as(C.POINTER(C.c_int32)), len(anchors),
                                                C.byref(out)), "bd_update_topology")
        return out.value

    def state(self):
        """(N,) uint8 node states: 0 cloth, 1 detached, 2 retired."""
        return self.read(BD_BUF_STATE)

    def set_node_state(self, state):
        self.write(BD_BUF_STATE, np.asarray(state, dtype=np.uint8))

    def rigidk(self):
        """(N, 2, 6) per detached node: external block, rotation stiffness (xx xy xz yy yz zz)."""
        return self.read(BD_BUF_RIGIDK).reshape(-1, 2, 6)

    def pieces(self):
        """Per piece: centroid (3), valid, E^-1 (6, 6), rigid step t (3), w (3)."""
        a = self.read(BD_BUF_PIECES).reshape(-1, 48)
        return {"centroid": a[:, :3], "valid": a[:, 3], "Einv": a[:, 4:40].reshape(-1, 6, 6),
                "t": a[:, 40:43], "w": a[:, 44:47]}

    def curve(self):
        """(N, 7): rotation w xyz, flag, arm r xyz."""
        return self.read(BD_BUF_CURVE).reshape(-1, 7)

    def autotune(self, reps=3):
        """Pick the fastest smoother kernel per level and family (bd_autotune): {level: (u, v)}."""
        nlev = len(self.stats()["level_nodes"])
        out = np.zeros(2 * nlev, dtype=np.int32)
        self._check(self.lib.bd_autotune(self.h, reps, out.ctypes.data_as(C.POINTER(C.c_int32))), "bd_autotune")
        return {k: (int(out[2 * k]), int(out[2 * k + 1])) for k in range(nlev)}

    def set_smoother(self, level, family, variant):
        """Pin the smoother kernel variant of one level's u-line (0) or v-line (1) passes."""
        self._check(self.lib.bd_set_smoother(self.h, level, family, variant), "bd_set_smoother")

    def sync(self):
        self._check(self.lib.bd_sync(self.h), "bd_sync")

    # ---- readback
    def read(self, buf, level=0):
        nbytes = self.lib.bd_buffer_bytes(self.h, buf, level)
        if nbytes == 0:
            raise RuntimeError(f"buffer {buf} level {level} is empty")
        dt = _DTYPES.get(buf, np.float64)
        out = np.empty(nbytes // np.dtype(dt).itemsize, dtype=dt)
        self._check(self.lib.bd_read(self.h, buf, level, out.ctypes.data, nbytes), "bd_read")
        return out

    def write(self, buf, data, level=0):
        nbytes = self.lib.bd_buffer_bytes(self.h, buf, level)
        dt = _DTYPES.get(buf, np.float64)
        a = np.ascontiguousarray(data, dtype=dt).ravel()
        if a.nbytes < nbytes:
            raise ValueError(f"need {nbytes} bytes for buffer {buf}, got {a.nbytes}")
        self._check(self.lib.bd_write(self.h, buf, level, a.ctypes.data, a.nbytes), "bd_write")

    def vec(self, buf, level=0):
        return self.read(buf, level).reshape(-1, 3)

    # ---- contact buffers
    def ext(self):
        """(N, 9) external terms: gradient xyz, Hessian xx xy xz yy yz zz."""
        return self.read(BD_BUF_EXT).reshape(-1, 9)

    def cmeta(self):
        """(N, 2, 4) bounds: own (normal, share) and reciprocal (direction, share)."""
        return self.read(BD_BUF_CMETA).reshape(-1, 2, 4)

    def pairs(self):
        """Self-contact: (N, K + 2) K candidate faces, reciprocal key, row mark. Plies: (N, K)
        candidate nodes of the other layers."""
        return self.read(BD_BUF_PAIRS).reshape(self.N, -1)

    def cells(self):
        return self.read(BD_BUF_CELLS)

    def items(self):
        """Self-contact: (2N, 8) face records (centroid now, radius, centroid at start, face id
        bits). Plies: (N,) uint32 node ids in cell order."""
        a = self.read(BD_BUF_ITEMS)
        return a.view(np.uint32) if self.contact_mode == "plies" else a.reshape(-1, 8)

    @property
    def x(self): return self.vec(BD_BUF_X)

    @property
    def v(self): return self.vec(BD_BUF_V)

    @property
    def dx(self): return self.vec(BD_BUF_DX)

    def stencil(self, level=0):
        return self.read(BD_BUF_STENCIL, level).reshape(-1, 25, 3, 3)

    def fixed(self, level=0):
        return self.read(BD_BUF_FIXED, level)

    def prolongation(self, level=0):
        return self.read(BD_BUF_PROLONG, level).reshape(-1, 4)

    def energy(self):
        e = self.read(BD_BUF_ENERGY)
        return e[:4], e[4:8]

    def line_alpha(self):
        """The step length the last line search chose."""
        return float(self.read(BD_BUF_ENERGY)[8])

    def stats(self):
        st = bd_stats()
        self._check(self.lib.bd_get_stats(self.h, C.byref(st)), "bd_get_stats")
        n = st.newton_iterations
        return {
            "newton_iterations": n,
            "pcg_iterations": list(st.pcg_iterations[:max(n, 1)]),
            "rel_residual": list(st.rel_residual[:max(n, 1)]),
            "alpha": list(st.alpha[:max(n, 1)]),
            "max_update": list(st.max_update[:max(n, 1)]),
            "rejections": st.rejections,
            "ruptures": st.ruptures,
            "pulled_out": st.pulled_out,
            "level_nodes": list(st.level_nodes[:st.level_count]),
            "poll_fallbacks": st.poll_fallbacks,
            "learned_fallbacks": st.learned_fallbacks,
            "warm_relres": list(st.warm_relres[:max(n, 1)]),
            "max_speed": st.max_speed,
        }

    # ---- learned coarse-space correction (bd_learn_*)
    def learn_info(self):
        lv, n, npad, cnt = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._check(self.lib.bd_learn_info(self.h, C.byref(lv), C.byref(n), C.byref(npad), C.byref(cnt)), "bd_learn_info")
        return {"level": lv.value, "nodes": n.value, "padded": npad.value, "samples": cnt.value}

    def learn_record(self, capacity):
        self._check(self.lib.bd_learn_record(self.h, int(capacity)), "bd_learn_record")

    def learn_train(self, epochs, lr=1e-3, batch=1, targets=3):
        losses = np.zeros(max(int(epochs), 1))
        self._check(self.lib.bd_learn_train(self.h, int(epochs), float(lr), int(batch), int(targets),
                                            losses.ctypes.data_as(C.POINTER(C.c_double))), "bd_learn_train")
        return losses[:int(epochs)]

    def set_learned(self, mode):
        """0 / False: plain MG-PCG; 1 / True: the learned warm start; 2: the warm start and the deflated
        correction (BD_LEARNED_*)."""
        self._check(self.lib.bd_set_learned(self.h, int(mode)), "bd_set_learned")

    def learn_weights(self):
        w = np.zeros(capi.BD_NET_PARAMS, np.float32)
        self._check(self.lib.bd_learn_get_weights(self.h, w.ctypes.data_as(C.POINTER(C.c_float)), w.size),
                    "bd_learn_get_weights")
        return w

    def set_learn_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(self.lib.bd_learn_set_weights(self.h, w.ctypes.data_as(C.POINTER(C.c_float)), w.size),
                    "bd_learn_set_weights")

    def learn_loss_grad(self, sample):
        loss = C.c_double()
        g = np.zeros(capi.BD_NET_PARAMS, np.float32)
        self._check(self.lib.bd_learn_loss_grad(self.h, int(sample), C.byref(loss),
                                                g.ctypes.data_as(C.POINTER(C.c_float)), g.size), "bd_learn_loss_grad")
        return loss.value, g

    def learn_sample(self, sample):
        """A recorded sample: targets (2, 3, Np_l), energies (2,), operator (78, Np_l), fixed flags (Np_l,)."""
        info = self.learn_info()
        npad = info["padded"]
        b = np.zeros((2, 3, npad), np.float32)
        e = np.zeros(2)
        op = np.zeros((78, npad), np.float32)
        fx = np.zeros(npad, np.uint8)
        self._check(self.lib.bd_learn_read_sample(self.h, int(sample), b.ctypes.data_as(C.POINTER(C.c_float)),
                                                  e.ctypes.data_as(C.POINTER(C.c_double)),
                                                  op.ctypes.data_as(C.POINTER(C.c_float)),
                                                  fx.ctypes.data_as(C.POINTER(C.c_uint8))), "bd_learn_read_sample")
        return b, e, op, fx

    def learn_fields(self):
        npad = self.learn_info()["padded"]
        z = np.zeros((capi.BD_NET_FIELDS, 3, npad), np.float32)
        self._check(self.lib.bd_learn_read_fields(self.h, z.ctypes.data_as(C.POINTER(C.c_float))), "bd_learn_read_fields")
        return z


class Batch:
    """Independent solvers (uneven patches, each its own Newton / MG-PCG system) advanced together.
    By default the form bd_batch_create picks: one graph launch per frame (every solver's captured
    substep a branch of the batch's graph), or, for solvers in BD_PCG_WORKGROUP mode, each Newton
    iteration's solves in one launch. With `threads` (an int, 0 for the library's default) the
    concurrent form: each solver's eager substep on its own stream from a pool of host threads."""

    def __init__(self, solvers, threads=None):
        self.lib = solvers[0].lib
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[g.h for g in self.solvers])
        out = C.c_void_p()
        if threads is None:
            rc = self.lib.bd_batch_create(arr, len(self.solvers), C.byref(out))
            what = "bd_batch_create"
        else:
            rc = self.lib.bd_batch_create_concurrent(arr, len(self.solvers), int(threads), C.byref(out))
            what = "bd_batch_create_concurrent"
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self.lib.bd_batch_last_error(None).decode()}")
        self.b = out

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self.lib.bd_batch_last_error(self.b).decode()}")

    def substep(self):
        self._check(self.lib.bd_batch_substep(self.b), "bd_batch_substep")

    def sync(self):
        self._check(self.lib.bd_batch_sync(self.b), "bd_batch_sync")

    def close(self):
        if getattr(self, "b", None):
            self.lib.bd_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        self.close()

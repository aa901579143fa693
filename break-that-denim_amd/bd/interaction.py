"""Picking and finite-patch grips over a GpuSolver (src/simulation/gpu/interaction.ts).

Up to BD_MAX_GRIPS_HELD grips (one per finger) hold disjoint patches, each with its own handle,
bounded force and pinch. Picking returns the material's identity (node), the picked point and the
connectivity window around it through a small readback. The GPU side is interact.hip behind the
bd_grip_* / bd_pick_* entry points; the patches are built here (grip.py).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .capi import BD_GRIP_SLOT_ENTRIES, BD_MAX_GRIPS_HELD, BD_PICK_MAX_WINDOW_RADIUS, bd_pick
from .grip import GRIP, GripPatch, PatchWindow, grip_patch, patch_window_radius

MAX_GRIPS = BD_MAX_GRIPS_HELD


@dataclass
class PickResult:
    """interaction.ts:22-31."""
    node: int
    position: tuple
    kind: str          # 'surface': the ray hit visible material; 'proximity': the nearest unoccluded crossing
    window: PatchWindow


@dataclass
class _Held:
    nodes: list
    count: int
    core: list
    releasing: bool
    stiffness: float


class Interaction:
    """interaction.ts:48-348 over the C-ABI."""

    def __init__(self, solver):
        self.solver = solver
        self.lib = solver.lib
        self.held = [None] * MAX_GRIPS
        self.owner = {}
        scene = solver.scene
        self.window_radius = patch_window_radius(scene)
        if self.window_radius > BD_PICK_MAX_WINDOW_RADIUS:
            raise ValueError(f"a grip patch spans {self.window_radius} lattice cells; the pick window covers "
                             f"{BD_PICK_MAX_WINDOW_RADIUS}")
        # strands over openings are picked within a quarter of the finer pitch (interaction.ts:88-90)
        self.strand_radius = 0.25 * min(min(L.material["pitchU"], L.material["pitchV"]) for L in scene.layers)

    def _result(self, out: bd_pick):
        if out.kind == 0:
            return None
        R = out.window_radius
        bits = np.array(out.window[:(2 * R + 1) ** 2], dtype=np.uint32)
        return PickResult(int(out.node), tuple(out.position), "surface" if out.kind == 1 else "proximity",
                          PatchWindow(R, bits))

    def pick_at(self, view_proj, px, py, width, height, radius_px, eye, direction):
        """pickAt (interaction.ts:114-144): the visible material under a drawing-buffer pixel, given the
        camera's eye and the unit ray through the pixel (view_proj column-major, as WebGPU's)."""
        vp = np.ascontiguousarray(view_proj, dtype=np.float32).reshape(16)
        e = np.ascontiguousarray(eye, dtype=np.float64)
        d = np.ascontiguousarray(direction, dtype=np.float64)
        out = bd_pick()
        f64 = C.POINTER(C.c_double)
        self.solver._check(self.lib.bd_pick_ray(self.solver.h, vp.ctypes.data_as(C.POINTER(C.c_float)), px, py, width,
                                                height, radius_px, self.strand_radius, self.window_radius,
                                                e.ctypes.data_as(f64), d.ctypes.data_as(f64), C.byref(out)),
                           "bd_pick_ray")
        return self._result(out)

    def nearest(self, p, radius):
        """nearest (interaction.ts:150-166): the crossing with material nearest to p within radius (m)."""
        q = np.ascontiguousarray(p, dtype=np.float64)
        out = bd_pick()
        self.solver._check(self.lib.bd_pick_near(self.solver.h, q.ctypes.data_as(C.POINTER(C.c_double)), radius,
                                                 self.window_radius, C.byref(out)), "bd_pick_near")
        return self._result(out)

    def free_slot(self):
        return next((s for s, h in enumerate(self.held) if h is None), -1)

    def holder(self, node):
        return self.owner.get(node, -1)

    def is_free(self, n):
        return n not in self.owner

    def attach(self, node, anchor, slot=0, window=None):
        """attach (interaction.ts:198-205): grip `slot` takes a pinch around `node`; False when no free
        material is left."""
        self.release(slot)
        patch = grip_patch(self.solver.scene, node, self.is_free, window)
        if patch is None:
            return False
        self.attach_patch(slot, node, patch, anchor, GRIP["maxLead"])
        return True

    def attach_patch(self, slot, picked, patch: GripPatch, anchor, max_lead):
        """attachPatch (interaction.ts:211-262)."""
        self.release(slot)
        count = min(len(patch.nodes), BD_GRIP_SLOT_ENTRIES)
        nodes = np.ascontiguousarray(patch.nodes[:count], dtype=np.int32)
        w = np.ascontiguousarray(patch.weights[:count], dtype=np.float32)
        sides = np.ascontiguousarray(patch.sides[:count], dtype=np.float32)
        a = np.ascontiguousarray(anchor, dtype=np.float64)
        self.solver._check(self.lib.bd_grip_attach(self.solver.h, slot, int(picked), nodes.ctypes.data_as(C.POINTER(C.c_int32)),
                                                   w.ctypes.data_as(C.POINTER(C.c_float)),
                                                   sides.ctypes.data_as(C.POINTER(C.c_float)), count, patch.stiffness,
                                                   max_lead, patch.pad_pitch, patch.slip,
                                                   a.ctypes.data_as(C.POINTER(C.c_double))), "bd_grip_attach")
        held_nodes = list(patch.nodes[:count])
        core = [n for n, wt in zip(patch.nodes[:count], patch.weights[:count]) if wt >= 0.5]
        # crossings on the grab line belong to this pinch (another grab may not take them) without being held
        held_nodes += list(patch.between)
        for n in held_nodes:
            self.owner[n] = slot
        self.held[slot] = _Held(held_nodes, count, core, False, patch.stiffness)

    def set_pinch(self, scale, slot=0):
        if self.held[slot] is not None:
            self.solver._check(self.lib.bd_grip_pinch(self.solver.h, slot, scale), "bd_grip_pinch")

    def set_hold(self, fraction, slot=0):
        if self.held[slot] is not None:
            self.solver._check(self.lib.bd_grip_hold(self.solver.h, slot, fraction), "bd_grip_hold")

    def set_target(self, target, slot=0):
        if self.held[slot] is not None:
            t = np.ascontiguousarray(target, dtype=np.float64)
            self.solver._check(self.lib.bd_grip_target(self.solver.h, slot, t.ctypes.data_as(C.POINTER(C.c_double))),
                               "bd_grip_target")

    def let_go(self, slot=0):
        """letGo (interaction.ts:296-306): still holding, but no longer an anchor of what it holds."""
        if self.held[slot] is not None:
            self.held[slot].releasing = True

    def release(self, slot=0):
        h = self.held[slot]
        if h is None:
            return
        for n in h.nodes:
            self.owner.pop(n, None)
        self.held[slot] = None
        self.solver._check(self.lib.bd_grip_release(self.solver.h, slot), "bd_grip_release")

    def release_all(self):
        for s in range(MAX_GRIPS):
            self.release(s)

    @property
    def active(self):
        return any(h is not None for h in self.held)

    @property
    def patch_nodes(self):
        return [n for h in self.held if h is not None for n in h.nodes]

    @property
    def core_nodes(self):
        """What the grips actually hold (full weight): the detached-piece anchors."""
        return [n for h in self.held if h is not None and not h.releasing for n in h.core]

    def encode_substep(self):
        """Once per substep before the Newton iterations (bd_substep runs it itself)."""
        self.solver._check(self.lib.bd_grip_lead(self.solver.h), "bd_grip_lead")

    def encode(self):
        """The grip terms into ext at the current iterate (bd_substep runs it each Newton iteration)."""
        self.solver._check(self.lib.bd_grip_terms(self.solver.h), "bd_grip_terms")

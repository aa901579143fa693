"""Grips and picking on the GPU (csrc/interact.hip) through the C-ABI, on a 40 x 40 coupon of 9 x 6
bundled denim (3.4 mm pitch) mounted by its collar, checked against what the construction fixes:

* at the grab the pinch's springs are slack (zero force, stiffness k w per entry);
* a handle far beyond the largest lead is clamped to it: the total grip force equals
  GRIP["maxForce"] (interact.wgsl gripLead);
* a closing pinch pulls each pad entry by (1 - s) pad pitches along the closing axis;
* an entry torn farther than the slip distance from the held patch slips out (weight 0);
* a handle lifted far away lifts the cloth with a force that never exceeds the bound;
* a pointer ray onto the cloth picks the surface at its intersection with the plane and the
  nearest corner, over a hole it picks the nearest unoccluded crossing on screen, and the nearest
  material to a point is the nearest crossing with a surviving span.
"""
import sys

import numpy as np
import pytest

from bd.capi import BD_BUF_SEGFLAGS, make_params
from bd.grip import GRIP, pinch_scale
from bd.scene import SEG_BROKEN, compile_scene, coupon

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sheet():
    return compile_scene([coupon(40, 40, bundle=[9, 6], collarRows=1)])


def setup(S, node):
    from bd.interaction import Interaction
    from bd.solver import GpuSolver
    gpu = GpuSolver(S, make_params(history=False))
    ia = Interaction(gpu)
    anchor = S.position[node].copy()
    assert ia.attach(node, anchor)
    return gpu, ia, anchor


def centre(S):
    L = S.layers[0]
    return 20 * L.nx + 20


def test_grip_terms_at_grab_and_lead_clamp():
    S = sheet()
    node = centre(S)
    gpu, ia, anchor = setup(S, node)
    h = ia.held[0]
    say(f"pinch of {h.count} entries, stiffness {h.stiffness:.1f} N/m per unit weight")
    gpu.encode_predict()
    ia.encode_substep()
    ia.encode()
    ext = gpu.ext()
    entries = np.array(h.nodes[:h.count])
    from bd.grip import grip_patch
    g = grip_patch(S, node)
    w = np.array(g.weights)
    assert np.abs(ext[:, :3]).max() < 1e-6 * GRIP["maxForce"], "slack at the grab"
    assert np.allclose(ext[entries, 3], h.stiffness * w, rtol=1e-6)
    assert np.allclose(ext[entries, 6], h.stiffness * w, rtol=1e-6) and np.allclose(ext[entries, 8], h.stiffness * w, rtol=1e-6)
    others = np.setdiff1d(np.arange(S.node_count), entries)
    assert (ext[others] == 0).all()
    # a handle 10 cm above: the lead clamp bounds the total force at maxForce
    ia.set_target(anchor + np.array([0, 0, 0.1]))
    ia.encode_substep()
    ia.encode()
    total = gpu.ext()[:, :3].sum(0)
    say(f"total grip gradient {total} (bound {GRIP['maxForce']} N)")
    assert np.linalg.norm(total) == pytest.approx(GRIP["maxForce"], rel=1e-4)
    assert total[2] < 0 and abs(total[0]) + abs(total[1]) < 1e-3 * GRIP["maxForce"]


def test_pinch_gathers_the_pads():
    S = sheet()
    node = centre(S)
    gpu, ia, anchor = setup(S, node)
    from bd.grip import grip_patch
    g = grip_patch(S, node)
    s = pinch_scale(1.0)
    ia.set_pinch(s)
    gpu.encode_predict()
    ia.encode_substep()
    ia.encode()
    ext = gpu.ext()
    off = S.position[g.nodes] - S.position[node]
    axis = (np.array(g.sides)[:, None] * np.array(g.weights)[:, None] * off).sum(0)
    axis /= np.linalg.norm(axis)
    want = (g.stiffness * np.array(g.weights) * np.array(g.sides) * (1 - s) * g.pad_pitch)[:, None] * axis
    err = np.abs(ext[g.nodes, :3] - want).max() / np.abs(want).max()
    say(f"pad pull at pinch {s:.2f}: {np.abs(want).max():.3f} N max, error {err:.2e}")
    assert err < 1e-4


def test_torn_entry_slips_out():
    S = sheet()
    node = centre(S)
    gpu, ia, anchor = setup(S, node)
    from bd.grip import grip_patch
    g = grip_patch(S, node)
    rim = [n for n, w in zip(g.nodes, g.weights) if w < 0.5]
    torn = rim[0]
    x = S.position.copy()
    x[torn] += np.array([0.0, 0.0, 0.02])  # 2 cm away: beyond the 8 mm slip distance
    gpu.set_state(x)
    gpu.encode_predict()
    ia.encode_substep()
    ia.encode()
    ext = gpu.ext()
    assert ext[torn, 3] == 0 and (ext[[n for n in g.nodes if n != torn], 3] > 0).all()


def test_lifted_handle_lifts_the_cloth_with_a_bounded_force():
    S = sheet()
    node = centre(S)
    gpu, ia, anchor = setup(S, node)
    z0 = S.position[node, 2]
    worst = 0.0
    for k in range(20):
        ia.set_target(anchor + np.array([0, 0, 0.005 * (k + 1)]))
        gpu.encode_substep()
        ia.encode()  # the terms at the committed state
        worst = max(worst, float(np.linalg.norm(gpu.ext()[:, :3].sum(0))))
    lift = gpu.x[node, 2] - z0
    say(f"lifted {lift * 1e3:.1f} mm; largest total grip force {worst:.2f} N (bound {GRIP['maxForce']})")
    assert np.isfinite(gpu.x).all() and lift > 0.01
    assert worst <= GRIP["maxForce"] * (1 + 1e-4)


def ortho(x0, y0, half, z_near=1.0, z_far=-1.0):
    """A column-major view-projection looking down -z: x, y in [x0 - half, x0 + half] map to NDC,
    z from z_near (depth 0) to z_far (depth 1), w = 1."""
    M = np.zeros((4, 4))
    M[0, 0] = 1 / half
    M[0, 3] = -x0 / half
    M[1, 1] = 1 / half
    M[1, 3] = -y0 / half
    M[2, 2] = -1 / (z_near - z_far)
    M[2, 3] = z_near / (z_near - z_far)
    M[3, 3] = 1
    return M.T.ravel()  # column-major


def test_pick_ray_hits_the_surface():
    from bd.interaction import Interaction
    from bd.solver import GpuSolver
    S = sheet()
    gpu = GpuSolver(S, make_params(history=False))
    ia = Interaction(gpu)
    L = S.layers[0]
    n = 17 * L.nx + 23
    p = S.position[n] + np.array([0.3e-3, -0.2e-3, 0.0])
    eye = p + np.array([0, 0, 0.5])
    r = ia.pick_at(ortho(p[0], p[1], 0.05), 400, 300, 800, 600, 12, eye, [0, 0, -1])
    say(f"ray pick: node {r.node} ({r.kind}) at {np.array(r.position)}")
    assert r.kind == "surface" and r.node == n
    assert np.allclose(r.position, [p[0], p[1], S.position[n, 2]], atol=2e-6)
    R = r.window.radius
    assert R == ia.window_radius and (r.window.bits == 3).all()


def test_pick_over_a_hole_takes_the_nearest_visible_crossing():
    from bd.interaction import Interaction
    from bd.solver import GpuSolver
    S = sheet()
    L = S.layers[0]
    c = 20 * L.nx + 20
    flags = S.seg_flags.copy()
    hole = [c + dj * L.nx + di for dj in (-1, 0, 1) for di in (-1, 0, 1)]
    for m in hole:  # every span touching the hole's nodes
        flags[2 * m] |= SEG_BROKEN
        flags[2 * m + 1] |= SEG_BROKEN
        flags[2 * (m - 1)] |= SEG_BROKEN
        flags[2 * (m - L.nx) + 1] |= SEG_BROKEN
    gpu = GpuSolver(S, make_params(history=False))
    gpu.write(BD_BUF_SEGFLAGS, flags)
    ia = Interaction(gpu)
    p = S.position[c]
    eye = p + np.array([0, 0, 0.5])
    half = 0.05
    # a square viewport (screen distance proportional to distance in the plane); 80 px = 10 mm: the
    # hole's rim (two pitches from its centre) is inside the cursor disc
    r = ia.pick_at(ortho(p[0], p[1], half), 400, 400, 800, 800, 80, eye, [0, 0, -1])
    say(f"pick over the hole: node {r.node} ({r.kind})")
    assert r.kind == "proximity"
    has = []
    for m in range(S.node_count):
        own = any((flags[2 * m + f] & 3) == 1 for f in (0, 1))
        inc = ((m % L.nx) > 0 and (flags[2 * (m - 1)] & 3) == 1) or ((m // L.nx) > 0 and (flags[2 * (m - L.nx) + 1] & 3) == 1)
        has.append(own or inc)
    d = np.where(has, np.hypot(*(S.position[:, :2] - p[:2]).T), np.inf)
    assert d[r.node] <= d.min() + 1e-9 and r.node not in hole


def test_pick_near_a_point():
    from bd.interaction import Interaction
    from bd.solver import GpuSolver
    S = sheet()
    gpu = GpuSolver(S, make_params(history=False))
    ia = Interaction(gpu)
    q = S.position[500] + np.array([1e-3, 0.5e-3, 2e-3])
    r = ia.nearest(q, 0.01)
    d = np.linalg.norm(S.position - q, axis=1)
    assert r.kind == "proximity" and r.node == int(np.argmin(d))
    assert ia.nearest(q + np.array([0, 0, 0.5]), 0.01) is None

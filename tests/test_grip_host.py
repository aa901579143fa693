"""Grip patches on the host (bd/grip.py, restating src/simulation/grip.ts), pinned by what the
construction fixes: the pinch schedule's end points and smoothness, the disc's smooth weights and
force bound, the pad split along the closing axis, the window search against scipy's Dijkstra, and a
slit's far lip left out of a pinch grown through the connectivity window."""
import math

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import dijkstra

from bd.grip import GRIP, PATCH_PATH, PatchWindow, grip_patch, hook_patch, patch_disc, patch_window_radius, \
    pinch_scale, reachable
from bd.scene import compile_scene, coupon


def test_pinch_schedule():
    assert pinch_scale(0.0) == 1.0 and pinch_scale(-1.0) == 1.0
    assert pinch_scale(GRIP["engageTime"]) == pytest.approx(GRIP["pinch"])
    assert pinch_scale(10.0) == pytest.approx(GRIP["pinch"])
    t = np.linspace(0, GRIP["engageTime"], 50)
    s = np.array([pinch_scale(v) for v in t])
    assert (np.diff(s) <= 0).all()
    # smoothstep: zero slope at both ends
    e = 1e-6
    assert abs(pinch_scale(e) - 1) < 1e-9 and abs(pinch_scale(GRIP["engageTime"] - e) - GRIP["pinch"]) < 1e-9


def scene(n=60):
    return compile_scene([coupon(n, n, collarRows=1)])


def test_patch_disc_weights_and_force_bound():
    S = scene()
    L = S.layers[0]
    node = 30 * L.nx + 30
    d = patch_disc(S, node)
    u0, v0 = S.rest_uv[node]
    R = GRIP["radius"]
    W = 0.0
    for n, w in zip(d.nodes, d.weights):
        r = math.hypot(S.rest_uv[n, 0] - u0, S.rest_uv[n, 1] - v0) / R
        assert r < 1 and not S.pinned[n]
        lin = 1.0 if r < 0.5 else 1 - (r - 0.5) / 0.5
        assert w == pytest.approx(lin * lin * (3 - 2 * lin))
        W += w
    # every free node inside 0.999 R with a weight above the cut-off is there
    inside = [m for m in range(S.node_count) if not S.pinned[m]
              and math.hypot(S.rest_uv[m, 0] - u0, S.rest_uv[m, 1] - v0) < 0.99 * R]
    assert set(inside) <= set(d.nodes)
    assert d.stiffness * GRIP["maxLead"] * W == pytest.approx(GRIP["maxForce"])
    assert d.slip == R


def test_pads_split_across_the_grab_line():
    S = scene()
    L = S.layers[0]
    node = 30 * L.nx + 30
    g = grip_patch(S, node)
    pitch = L.material["pitchU"]
    assert g.pad_pitch == pitch
    c = S.rest_uv[node, 0]
    for n, side in zip(g.nodes, g.sides):
        assert side == np.sign(S.rest_uv[n, 0] - c) and abs(S.rest_uv[n, 0] - c) >= 0.5 * pitch
    for n in g.between:
        assert abs(S.rest_uv[n, 0] - c) < 0.5 * pitch
    W = sum(g.weights)
    assert g.stiffness * GRIP["maxLead"] * W == pytest.approx(GRIP["maxForce"])
    # at the lattice's left edge the pads along u are lopsided: the pinch closes along v instead
    edge = 30 * L.nx + 1
    e = grip_patch(S, edge)
    assert e.pad_pitch == L.material["pitchV"]


def window_of(S, node, R, broken=()):
    """The connectivity window around node (interact.wgsl pickWindow) from the scene's flags."""
    L = S.layers[0]
    i0, j0 = node % L.nx, node // L.nx
    side = 2 * R + 1
    bits = np.zeros(side * side, np.uint32)
    for k in range(side * side):
        i, j = i0 + k % side - R, j0 + k // side - R
        if not (0 <= i < L.nx and 0 <= j < L.ny):
            bits[k] = 8
            continue
        m = j * L.nx + i
        for f in (0, 1):
            if (S.seg_flags[2 * m + f] & 3) == 1 and (m, f) not in broken:
                bits[k] |= 1 << f
    return PatchWindow(R, bits)


@pytest.mark.parametrize("seed", [0, 1])
def test_reachable_against_dijkstra(seed):
    rng = np.random.default_rng(seed)
    R = 7
    side = 2 * R + 1
    bits = rng.integers(0, 4, side * side).astype(np.uint32)
    bits[rng.random(side * side) < 0.1] |= 4
    bits[R * side + R] &= 3
    pu, pv = 0.4e-3, 0.6e-3
    maxp = 3.1e-3
    got = reachable(PatchWindow(R, bits), pu, pv, maxp)
    rows, cols, w = [], [], []
    for k in range(side * side):
        i, j = k % side, k // side
        if bits[k] & 12:
            continue
        if i + 1 < side and bits[k] & 1 and not bits[k + 1] & 12:
            rows += [k]; cols += [k + 1]; w += [pu]
        if j + 1 < side and bits[k] & 2 and not bits[k + side] & 12:
            rows += [k]; cols += [k + side]; w += [pv]
    g = coo_matrix((w, (rows, cols)), shape=(side * side, side * side))
    dist = dijkstra(g, directed=False, indices=R * side + R)
    want = (dist <= maxp).astype(np.uint8)
    assert np.array_equal(got, want)


def test_slit_far_lip_is_not_pinched():
    """A slit (a row of broken v spans) beside the grab point: the far lip is within the disc's radius
    but only reachable around the slit's end, farther than the patch path allows."""
    S = scene()
    L = S.layers[0]
    node = 30 * L.nx + 30
    R = patch_window_radius(S)
    # break the v spans between rows 31 and 32 over the whole window width
    broken = {((31) * L.nx + i, 1) for i in range(30 - R - 2, 30 + R + 3)}
    win = window_of(S, node, R, broken)
    g = grip_patch(S, node, window=win)
    rows = {n // L.nx for n in g.nodes + g.between}
    assert max(rows) <= 31 and min(rows) < 30
    full = grip_patch(S, node)
    assert max(n // L.nx for n in full.nodes) > 32


def test_hook_patch_takes_the_larger_weight():
    S = scene()
    L = S.layers[0]
    claws = [30 * L.nx + 28, 30 * L.nx + 32]
    h = hook_patch(S, claws, 0.002, 60.0, 0.004)
    a = patch_disc(S, claws[0], R=0.002)
    b = patch_disc(S, claws[1], R=0.002)
    wa = dict(zip(a.nodes, a.weights))
    wb = dict(zip(b.nodes, b.weights))
    for n, w in zip(h.nodes, h.weights):
        assert w == max(wa.get(n, 0), wb.get(n, 0))
    assert h.stiffness * 0.004 * sum(h.weights) == pytest.approx(60.0) and h.slip == 0

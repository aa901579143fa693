"""Grip patches on the host (src/simulation/grip.ts): the smoothly weighted material disc a fingertip
pinch holds, split into two pads either side of the grab line, claws hooked into the weave, the
pinch's closing schedule and the connectivity search the patch grows through.

A grip is a quadratic spring from each patch node to the handle (exact for Newton); its force is
bounded by limiting how far the handle may lead the gripped material (the GPU's lead clamp), so the
total force stays below GRIP["maxForce"].
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# grip.ts:8-35
GRIP = {
    "radius": 0.008,       # material radius of the gripped patch (m): a fingertip pinch
    "maxForce": 120.0,     # total grip force bound (N)
    "maxLead": 0.006,      # largest lead of the handle over the patch centroid (m)
    "pinch": 0.3,          # spread kept by a closed pinch
    "pinchMode": "pads",   # 'pads' | 'radial'
    "pinchAxis": 0,        # material axis the fingertips close along when the cloth allows (0: u)
    "engageTime": 0.15,    # time over which the fingers close and lift the cloth (s)
}
# longest path through surviving spans, in patch radii, by which a patch member may be reached
PATCH_PATH = 1.5


def pinch_scale(t):
    """pinchScale (grip.ts:37-40): the fraction of the grab-time spread kept t seconds after the grab."""
    e = min(1.0, max(0.0, t / GRIP["engageTime"]))
    return 1 - (1 - GRIP["pinch"]) * e * e * (3 - 2 * e)


@dataclass
class GripPatch:
    """grip.ts:42-61."""
    nodes: list
    weights: list
    sides: list              # pad of each entry (-1, +1), 0 in a radial pinch
    pad_pitch: float         # lattice pitch along the closing axis (m)
    between: list            # crossings between the pads: inside the pinch, not held
    stiffness: float         # per unit weight (N/m): the summed force reaches maxForce at the largest lead
    slip: float              # an entry displaced farther than this from the held patch slips out (0: never)


@dataclass
class PatchWindow:
    """Connectivity around a picked node (grip.ts:63-71): per offset in [-radius, radius]^2, row-major
    from (-radius, -radius): bit 0 / 1 its own span along u / v survives, bit 2 not cloth, bit 3
    outside the lattice."""
    radius: int
    bits: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))


def patch_window_radius(scene):
    """patchWindowRadius (grip.ts:76-80): lattice cells either side of a picked node a window covers."""
    pitch = min(min(L.material["pitchU"], L.material["pitchV"]) for L in scene.layers)
    return int(math.ceil(GRIP["radius"] / pitch))


def _layer_of(scene, node):
    for L in scene.layers:
        if L.node_offset <= node < L.node_offset + L.node_count:
            return L
    return None


def reachable(win: PatchWindow, pitch_u, pitch_v, max_path):
    """reachable (grip.ts:196-222): window cells reached from its centre through surviving spans
    within path length max_path (m), over cloth crossings only (Dijkstra)."""
    side = 2 * win.radius + 1
    count = side * side
    dist = np.full(count, np.inf)
    out = np.zeros(count, np.uint8)
    bits = np.asarray(win.bits)
    blocked = lambda q: (int(bits[q]) & 12) != 0
    centre = win.radius * side + win.radius
    if blocked(centre):
        return out
    dist[centre] = 0.0
    while True:
        k = -1
        for q in range(count):
            if not out[q] and dist[q] <= max_path and (k < 0 or dist[q] < dist[k]):
                k = q
        if k < 0:
            return out
        out[k] = 1
        i, j = k % side, k // side

        def relax(q, length):
            if not blocked(q) and not out[q]:
                dist[q] = min(dist[q], dist[k] + length)

        if i + 1 < side and bits[k] & 1:
            relax(k + 1, pitch_u)
        if i > 0 and bits[k - 1] & 1:
            relax(k - 1, pitch_u)
        if j + 1 < side and bits[k] & 2:
            relax(k + side, pitch_v)
        if j > 0 and bits[k - side] & 2:
            relax(k - side, pitch_v)


def patch_disc(scene, node, available=lambda n: True, window: PatchWindow | None = None, R=None):
    """patchDisc (grip.ts:157-193): the smoothly weighted disc of material radius R around `node`
    (same layer, free nodes only), reached through surviving spans when a window is given."""
    R = GRIP["radius"] if R is None else R
    L = _layer_of(scene, node)
    if L is None:
        return None
    u0, v0 = scene.rest_uv[node]
    m = L.material
    reach = reachable(window, m["pitchU"], m["pitchV"], PATCH_PATH * R) if window is not None else None
    ri = window.radius if window is not None else int(math.ceil(R / m["pitchU"]))
    rj = window.radius if window is not None else int(math.ceil(R / m["pitchV"]))
    side = 2 * ri + 1
    local = node - L.node_offset
    i0, j0 = local % L.nx, local // L.nx
    nodes, weights = [], []
    W = 0.0
    for j in range(max(0, j0 - rj), min(L.ny - 1, j0 + rj) + 1):
        for i in range(max(0, i0 - ri), min(L.nx - 1, i0 + ri) + 1):
            n = L.node_offset + j * L.nx + i
            if scene.pinned[n] or not available(n):
                continue
            if reach is not None and not reach[(j - j0 + rj) * side + (i - i0 + ri)]:
                continue
            d = math.hypot(scene.rest_uv[n, 0] - u0, scene.rest_uv[n, 1] - v0) / R
            if d >= 1:
                continue
            # smooth falloff: full weight in the core, tapering to zero at the rim
            w = 1.0 if d < 0.5 else 1 - (d - 0.5) / 0.5
            ws = w * w * (3 - 2 * w)
            if ws <= 1e-3:
                continue
            nodes.append(n)
            weights.append(ws)
            W += ws
    if not nodes:
        return None
    return GripPatch(nodes, weights, [0] * len(nodes), 0.0, [], GRIP["maxForce"] / (GRIP["maxLead"] * W), R)


def _split_pads(scene, node, disc: GripPatch, axis):
    """splitPads (grip.ts:106-128): the disc split into two pads either side of the grab line across
    `axis`, or None when lopsided (neither pad under a quarter of the other)."""
    L = _layer_of(scene, node)
    pitch = L.material["pitchU"] if axis == 0 else L.material["pitchV"]
    c = scene.rest_uv[node, axis]
    out = GripPatch([], [], [], pitch, [], 0.0, disc.slip)
    side = [0.0, 0.0]
    for n, w in zip(disc.nodes, disc.weights):
        u = scene.rest_uv[n, axis] - c
        if abs(u) < 0.5 * pitch:
            out.between.append(n)
            continue
        out.nodes.append(n)
        out.weights.append(w)
        out.sides.append(1 if u > 0 else -1)
        side[1 if u > 0 else 0] += w
    if min(side) < 0.25 * max(side) or side[0] == 0:
        return None
    out.stiffness = GRIP["maxForce"] / (GRIP["maxLead"] * (side[0] + side[1]))
    return out


def grip_patch(scene, node, available=lambda n: True, window: PatchWindow | None = None):
    """gripPatch (grip.ts:93-104): the pinch a fingertip grab holds, as two pads along the preferred
    axis when the cloth reaches both sides, else along the other axis, else the radial disc."""
    disc = patch_disc(scene, node, available, window)
    if disc is None or GRIP["pinchMode"] == "radial":
        return disc
    for axis in (GRIP["pinchAxis"], 1 - GRIP["pinchAxis"]):
        pads = _split_pads(scene, node, disc, axis)
        if pads is not None:
            return pads
    return disc


def hook_patch(scene, claws, radius, max_force, max_lead, available=lambda n: True, window=None):
    """hookPatch (grip.ts:139-156): claws hooked into the cloth at several crossings, the union of a
    disc of `radius` around each (a node in two discs takes the larger weight); no pinch, no slip."""
    weight = {}
    for c in claws:
        disc = patch_disc(scene, c, available, window, radius)
        if disc is None:
            continue
        for n, w in zip(disc.nodes, disc.weights):
            weight[n] = max(weight.get(n, 0.0), w)
    if not weight:
        return None
    nodes = list(weight)
    weights = [weight[n] for n in nodes]
    return GripPatch(nodes, weights, [0] * len(nodes), 0.0, [], max_force / (max_lead * sum(weights)), 0.0)

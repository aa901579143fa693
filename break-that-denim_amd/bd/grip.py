"""Grip patches on the host: which crossings a grab holds, with what weight, and how hard.

Interface after the reference's grip module (src/simulation/grip.ts: pinchScale, patchDisc,
gripPatch, hookPatch, reachable, patchWindowRadius, the GRIP constants); the computation here is
this build's own: every patch is evaluated on the lattice window around the grabbed crossing as
numpy arrays (rest-frame distances, a smoothstep falloff, pad membership by the sign of the offset
along the closing axis), and the reach through surviving spans is a heap-ordered shortest-path
search over the window's span graph.

A grip is a quadratic spring from each patch node to the handle (exact for Newton). Its force is
bounded by how far the handle may lead the held material (the GPU's lead clamp): the stiffness per
unit weight is chosen so that the whole patch pulls GRIP["maxForce"] at the largest lead.
"""
from __future__ import annotations

import heapq
import math
from dataclasses import dataclass, field

import numpy as np

# The reference's grip constants (grip.ts:8-35), kept as the interface's parameters.
GRIP = {
    "radius": 0.008,
    "maxForce": 120.0,
    "maxLead": 0.006,
    "pinch": 0.3,
    "pinchMode": "pads",
    "pinchAxis": 0,
    "engageTime": 0.15,
}
# a patch member must be reachable through surviving spans within this many patch radii
PATCH_PATH = 1.5

# window bits (the GPU's pick window layout)
BIT_SPAN_U, BIT_SPAN_V, BIT_NOT_CLOTH, BIT_OUTSIDE = 1, 2, 4, 8


def pinch_scale(t):
    """Share of the grab-time spread still kept t seconds after the grab: a smoothstep from 1 down to
    GRIP["pinch"] over GRIP["engageTime"]."""
    e = float(np.clip(t / GRIP["engageTime"], 0.0, 1.0))
    s = e * e * (3.0 - 2.0 * e)
    return 1.0 + (GRIP["pinch"] - 1.0) * s


@dataclass
class GripPatch:
    """A held patch: entries (nodes, weights, pad sides -1 / +1 or 0 for a radial pinch), the pitch
    along the closing axis, the crossings between the pads (inside the pinch, not held), the spring
    stiffness per unit weight, and the displacement beyond which an entry slips out (0: never)."""
    nodes: list
    weights: list
    sides: list
    pad_pitch: float
    between: list
    stiffness: float
    slip: float


@dataclass
class PatchWindow:
    """Lattice connectivity around a grabbed node: per offset in [-radius, radius]^2 (row-major from
    (-radius, -radius)) the window bits BIT_SPAN_U / BIT_SPAN_V (its own span survives),
    BIT_NOT_CLOTH and BIT_OUTSIDE."""
    radius: int
    bits: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))


def patch_window_radius(scene):
    """Lattice cells either side of the grabbed node that a patch of GRIP["radius"] can reach."""
    pitch = min(min(L.material["pitchU"], L.material["pitchV"]) for L in scene.layers)
    return int(math.ceil(GRIP["radius"] / pitch))


def _layer(scene, node):
    for L in scene.layers:
        if L.node_offset <= node < L.node_offset + L.node_count:
            return L
    return None


def reachable(win: PatchWindow, pitch_u, pitch_v, max_path):
    """Window cells whose shortest path from the centre through surviving spans, over cloth crossings
    only, is at most max_path (m); uint8 flags per cell."""
    side = 2 * win.radius + 1
    bits = np.asarray(win.bits, dtype=np.uint32)
    out = np.zeros(side * side, np.uint8)
    start = win.radius * (side + 1)
    if bits[start] & (BIT_NOT_CLOTH | BIT_OUTSIDE):
        return out
    best = {start: 0.0}
    heap = [(0.0, start)]
    while heap:
        d, k = heapq.heappop(heap)
        if out[k] or d > best.get(k, math.inf):
            continue
        out[k] = 1
        i, j = k % side, k // side
        # the four lattice moves and the span each crosses (owned by the lower / left end)
        for q, owner, bit, step, ok in ((k + 1, k, BIT_SPAN_U, pitch_u, i + 1 < side),
                                        (k - 1, k - 1, BIT_SPAN_U, pitch_u, i > 0),
                                        (k + side, k, BIT_SPAN_V, pitch_v, j + 1 < side),
                                        (k - side, k - side, BIT_SPAN_V, pitch_v, j > 0)):
            if not ok or out[q] or not (bits[owner] & bit) or bits[q] & (BIT_NOT_CLOTH | BIT_OUTSIDE):
                continue
            nd = d + step
            if nd <= max_path and nd < best.get(q, math.inf):
                best[q] = nd
                heapq.heappush(heap, (nd, q))
    return out


def _disc_arrays(scene, node, available, window, R):
    """The disc's candidate nodes and smoothstep weights as arrays (same layer, free, available,
    reachable when a window is given); empty arrays when nothing qualifies."""
    L = _layer(scene, node)
    if L is None:
        return None, None, None
    m = L.material
    if window is not None:
        ri = rj = window.radius
        reach = reachable(window, m["pitchU"], m["pitchV"], PATCH_PATH * R).astype(bool)
    else:
        ri, rj = int(math.ceil(R / m["pitchU"])), int(math.ceil(R / m["pitchV"]))
        reach = None
    local = node - L.node_offset
    i0, j0 = local % L.nx, local // L.nx
    di, dj = np.meshgrid(np.arange(-ri, ri + 1), np.arange(-rj, rj + 1), indexing="xy")
    di, dj = di.ravel(), dj.ravel()
    inside = (i0 + di >= 0) & (i0 + di < L.nx) & (j0 + dj >= 0) & (j0 + dj < L.ny)
    if reach is not None:
        inside &= reach
    n = L.node_offset + (j0 + dj) * L.nx + (i0 + di)
    n = n[inside]
    keep = (scene.pinned[n] == 0) & np.array([bool(available(int(q))) for q in n], dtype=bool)
    n = n[keep]
    rel = np.hypot(*(scene.rest_uv[n] - scene.rest_uv[node]).T) / R
    # full weight inside half the radius, a smoothstep falloff to the rim
    ramp = np.clip(2.0 * (1.0 - rel), 0.0, 1.0)
    w = ramp * ramp * (3.0 - 2.0 * ramp)
    sel = (rel < 1.0) & (w > 1e-3)
    return L, n[sel], w[sel]


def patch_disc(scene, node, available=lambda n: True, window: PatchWindow | None = None, R=None):
    """The smoothly weighted material disc of radius R (default GRIP["radius"]) around `node`, as a
    radial patch; None when it holds nothing."""
    R = GRIP["radius"] if R is None else R
    L, n, w = _disc_arrays(scene, node, available, window, R)
    if L is None or len(n) == 0:
        return None
    return GripPatch(n.tolist(), w.tolist(), [0] * len(n), 0.0, [], GRIP["maxForce"] / (GRIP["maxLead"] * float(w.sum())), R)


def _pads(scene, node, disc: GripPatch, axis):
    """The disc as two pads either side of the grab line across material axis `axis` (the crossings
    within half a pitch of the line lie between them); None when one pad carries less than a quarter
    of the other's weight."""
    L = _layer(scene, node)
    pitch = L.material["pitchU"] if axis == 0 else L.material["pitchV"]
    n = np.asarray(disc.nodes)
    w = np.asarray(disc.weights)
    off = scene.rest_uv[n, axis] - scene.rest_uv[node, axis]
    mid = np.abs(off) < 0.5 * pitch
    side = np.where(off > 0, 1, -1)
    wl, wr = float(w[~mid & (side < 0)].sum()), float(w[~mid & (side > 0)].sum())
    if wl == 0.0 or min(wl, wr) < 0.25 * max(wl, wr):
        return None
    held = ~mid
    return GripPatch(n[held].tolist(), w[held].tolist(), side[held].tolist(), pitch, n[mid].tolist(),
                     GRIP["maxForce"] / (GRIP["maxLead"] * (wl + wr)), disc.slip)


def grip_patch(scene, node, available=lambda n: True, window: PatchWindow | None = None):
    """A fingertip grab: two pads along the preferred axis when the cloth reaches both sides of the
    grab line, else along the other axis, else the radial disc."""
    disc = patch_disc(scene, node, available, window)
    if disc is None or GRIP["pinchMode"] == "radial":
        return disc
    first = GRIP["pinchAxis"]
    return _pads(scene, node, disc, first) or _pads(scene, node, disc, 1 - first) or disc


def hook_patch(scene, claws, radius, max_force, max_lead, available=lambda n: True, window=None):
    """Claws hooked into the cloth at several crossings: the union of a disc of `radius` around each
    (a node in several discs keeps its largest weight); no pinch and no slip."""
    weight = {}
    for c in claws:
        disc = patch_disc(scene, c, available, window, radius)
        if disc is None:
            continue
        for n, w in zip(disc.nodes, disc.weights):
            if w > weight.get(n, 0.0):
                weight[n] = w
    if not weight:
        return None
    nodes = sorted(weight)
    weights = [weight[n] for n in nodes]
    return GripPatch(nodes, weights, [0] * len(nodes), 0.0, [], max_force / (max_lead * sum(weights)), 0.0)

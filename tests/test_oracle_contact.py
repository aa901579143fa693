"""Pins the FP32 restatement of the reference's self-contact (oracle/contact_oracle.c).

The reference has no CPU counterpart of selfcontact.wgsl, so each piece is checked against what
the mathematics fixes: closest points by dense sampling, segment/triangle crossings by a linear
solve, the hash broad phase against a scan of every face, the K nearest faces against an
independent numpy search, the penalty force as the derivative of its cubic energy, momentum
balance of the corner reactions, and the bound behaviour of tests/gpu/interaction.ts:185-248
(node and face driven through each other stay 0.3 mm apart; a fast node is stopped by the face its
path crosses; a lip across a tear inside the exclusion window is a contact, the linked cloth is not).
"""
import numpy as np
import pytest

import oracle as O
from bd.configs import config_fold, sheet_scene
from bd.material import linear_sheet_material

K = 4
NONE = 0xFFFFFFFF


def test_closest_tri_is_the_sampled_minimum():
    rng = np.random.default_rng(11)
    g = np.linspace(0.0, 1.0, 161)
    U, V = np.meshgrid(g, g)
    keep = U + V <= 1.0
    U, V = U[keep], V[keep]
    for _ in range(60):
        a, b, c = (rng.standard_normal(3) * 2e-3 for _ in range(3))
        p = rng.standard_normal(3) * 2e-3
        w, d2 = O.closest_tri(p, a, b, c)
        assert abs(w.sum() - 1) < 1e-5 and (w >= -1e-6).all()
        q = w[0] * a + w[1] * b + w[2] * c
        assert abs(np.dot(p - q, p - q) - d2) <= 1e-5 * max(d2, 1e-12) + 1e-13
        pts = a[None] + U[:, None] * (b - a)[None] + V[:, None] * (c - a)[None]
        dmin = ((pts - p) ** 2).sum(1).min()
        # the sampled minimum is never below the exact one, and close above it
        assert d2 <= dmin * (1 + 1e-5) + 1e-14
        assert dmin - d2 <= 2 * np.sqrt(dmin) * np.linalg.norm(b - a + c - a) / 160 + 1e-12


def test_face_hit_over_a_flat_cell():
    p00, p10, p01, p11 = np.array([0, 0, 0.]), np.array([1e-3, 0, 0.]), np.array([0, 1e-3, 0.]), np.array([1e-3, 1e-3, 0.])
    for (x, y, z) in [(3e-4, 2e-4, 5e-4), (8e-4, 6e-4, -2e-4), (5e-4, 5e-4, 1e-3)]:
        w, q, d = O.face_hit([x, y, z], p00, p10, p01, p11)
        assert abs(d - abs(z)) < 1e-9
        assert np.allclose(q, [x, y, 0], atol=1e-9)
        assert abs(w.sum() - 1) < 1e-6
    # beyond an edge: the distance to that edge
    w, q, d = O.face_hit([2e-3, 5e-4, 1e-3], p00, p10, p01, p11)
    assert abs(d - np.hypot(1e-3, 1e-3)) < 1e-9 and w[0] == 0 and w[2] == 0


def test_seg_tri_matches_a_linear_solve():
    rng = np.random.default_rng(5)
    hits = 0
    for _ in range(400):
        a, b, c = (rng.standard_normal(3) for _ in range(3))
        o = rng.standard_normal(3)
        d = rng.standard_normal(3) * 2
        t = O.seg_tri(o, d, a, b, c)
        M = np.stack([b - a, c - a, -d], 1)
        u, v, tt = np.linalg.solve(M, o - a)
        inside = u >= 0 and v >= 0 and u + v <= 1 and 0 <= tt <= 1
        margin = min(abs(u), abs(v), abs(1 - u - v), abs(tt), abs(1 - tt))
        if margin < 1e-4:
            continue
        if inside:
            hits += 1
            assert abs(t - tt) < 1e-4
        else:
            assert t == 2.0
    assert hits > 10


def test_seg_dist_closed_form():
    rng = np.random.default_rng(2)
    for _ in range(100):
        q, a, b = (rng.standard_normal(3) for _ in range(3))
        e = b - a
        s = np.clip(np.dot(q - a, e) / np.dot(e, e), 0, 1)
        assert abs(O.seg_dist(q, a, b) - np.linalg.norm(q - a - s * e)) < 1e-5


def test_hash_cell_formula():
    for c in [(0, 0, 0), (1, 2, 3), (-1, -5, 7), (100, -200, 300)]:
        u = [v & 0xFFFFFFFF for v in c]
        h = ((u[0] * 73856093) & 0xFFFFFFFF) ^ ((u[1] * 19349663) & 0xFFFFFFFF) ^ ((u[2] * 83492791) & 0xFFFFFFFF)
        assert O.hash_cell(*c) == h & ((1 << 17) - 1)


def _fold_state(push=2.5e-3, noise=1e-4, seed=3):
    cfg = config_fold()
    S = cfg["scene"]
    x = S.position.astype(np.float32)
    dx = np.zeros_like(x)
    dx[x[:, 2] > 1.5e-3, 2] = -push
    dx += (np.random.default_rng(seed).standard_normal(dx.shape) * noise).astype(np.float32)
    return S, x, dx


def test_hashed_detection_equals_the_scan_of_every_face():
    S, x, dx = _fold_state()
    T = O.ContactTwin(S, x, dx)
    cells, items, pairs, cmeta = T.detect()
    pb, cb = T.detect_brute()
    assert np.array_equal(pairs[:, :K], pb[:, :K])
    assert np.array_equal(cmeta.view(np.uint32), cb.view(np.uint32))
    # the table: starts are the exclusive scan of the per-cell counts, one record per cell a face is in
    T2 = 1 << 17
    starts, counts = cells[:T2].astype(np.int64), cells[T2:].astype(np.int64)
    assert np.array_equal(starts, np.concatenate([[0], np.cumsum(counts)[:-1]]))
    nvalid = sum(T.face_valid(f) for f in range(S.node_count))
    assert nvalid <= counts.sum() <= 2 * nvalid
    faces = items[: counts.sum(), 7].view(np.uint32)
    assert set(faces.tolist()) == {f for f in range(S.node_count) if T.face_valid(f)}
    assert (pairs[:, :K] != NONE).any() and (cmeta[:, 0, 3] < 1e8).any() and (cmeta[:, 1, 3] < 1e8).any()


def test_k_nearest_faces_match_an_independent_search():
    S, x, dx = _fold_state()
    T = O.ContactTwin(S, x, dx)
    _, _, pairs, _ = T.detect()
    L = S.layers[0]
    nx, ny = L.nx, L.ny
    p = x + dx
    valid = [f for f in range(S.node_count) if T.face_valid(f)]
    rng = np.random.default_rng(0)
    for n in rng.choice(S.node_count, 60, replace=False):
        ni, nj = n % nx, n // nx
        cand = []
        for f in valid:
            fi, fj = f % nx, f // nx
            if -3 <= fi - ni <= 2 and -3 <= fj - nj <= 2:
                continue  # own intact cloth (no damage in this scene)
            c = [f, f + 1, f + nx, f + nx + 1]
            w, q, d = O.face_hit(p[n], *(p[k] for k in c))
            if d < 3e-3:
                cand.append((float(d), f))
        cand.sort()
        want = [f for _, f in cand[:K]] + [NONE] * (K - min(K, len(cand)))
        assert list(pairs[n, :K]) == want, n


def test_penalty_is_the_derivative_of_the_cubic_energy():
    # one node over the centre of a flat cell, inside the contact distance
    mat = linear_sheet_material(3.4e-3, 2e4, 1.6, 1e-5, 0.35, 8e3)
    S = sheet_scene(32, 3.4e-3, mat, perturbation=0.0, pin_top=False)
    x = S.position.astype(np.float32)
    x[:, 1] += 0.3  # the sheet far away
    nx = 32
    f = 10 * nx + 10
    node = 25 * nx + 25
    for q, c in enumerate([f, f + 1, f + nx, f + nx + 1]):
        x[c] = [((q % 2) - 0.5) * 3.4e-3, ((q // 2) - 0.5) * 3.4e-3, 0.0]
    kc, delta = 500.0, 1e-3
    for depth in (0.2e-3, 0.5e-3, 0.8e-3):
        x[node] = [0.0, 0.0, delta - depth]
        dx = np.zeros_like(x)
        T = O.ContactTwin(S, x, dx)
        _, _, pairs, cmeta = T.detect()
        assert f in pairs[node, :K]
        acc, ext = T.terms(pairs, cmeta.copy(), 1 / 120, 9.81)
        E = lambda s: kc * s ** 3 / (3 * delta)  # noqa: E731
        h = 1e-7
        dEdz = -(E(depth + h) - E(depth - h)) / (2 * h)   # z up reduces the depth
        d2Edz2 = (E(depth + h) - 2 * E(depth) + E(depth - h)) / h ** 2
        assert abs(ext[node, 2] - dEdz) < 1e-6 + 1e-5 * abs(dEdz)
        # the node takes twice the exact stiffness (the pair's closure is split between node and face)
        assert abs(ext[node, 8] - 2 * d2Edz2) < 2e-3 + 1e-4 * d2Edz2
        # the corners take the reaction: momentum balances to the fixed-point resolution
        assert np.abs(acc[:, :3].astype(np.int64).sum(0)).max() <= 8


def test_friction_is_coulomb_above_the_regularisation_speed():
    mat = linear_sheet_material(3.4e-3, 2e4, 1.6, 1e-5, 0.35, 8e3)
    S = sheet_scene(32, 3.4e-3, mat, perturbation=0.0, pin_top=False)
    x = S.position.astype(np.float32)
    x[:, 1] += 0.3
    nx = 32
    f = 10 * nx + 10
    node = 25 * nx + 25
    for q, c in enumerate([f, f + 1, f + nx, f + nx + 1]):
        x[c] = [((q % 2) - 0.5) * 3.4e-3, ((q // 2) - 0.5) * 3.4e-3, 0.0]
    x[node] = [0.0, 0.0, 0.5e-3]
    dx = np.zeros_like(x)
    dx[node] = [0.2e-3, 0.0, 0.0]    # 24 mm/s tangential step at h = 1/120, above 2 mm/s
    T = O.ContactTwin(S, x, dx)
    _, _, pairs, cmeta = T.detect()
    acc, ext = T.terms(pairs, cmeta.copy(), 1 / 120, 9.81)
    lam = 500.0 * (0.5e-3) ** 2 / 1e-3
    # the node sits at 0.5 mm + 0 over the face after its step: depth 0.5 mm
    assert abs(-ext[node, 2] - lam) < 1e-5
    assert abs(ext[node, 0] - 0.45 * lam) < 1e-5


def _probe_scene():
    mat = linear_sheet_material(3.42e-3, 2e4, 1.6, 1e-5, 0.35, 8e3)
    S = sheet_scene(64, 3.42e-3, mat, perturbation=0.0, pin_top=False)
    x = S.position.astype(np.float32)
    x[:, 1] += 0.3
    return S, x


def _apply(S, x, placed, delta, velocity=None, flags=None, h=1 / 120):
    """detect at the predicted iterate, then one full Newton update made feasible (newtonApply)."""
    x = x.copy()
    for n, p in placed.items():
        x[n] = p
    dx = np.zeros_like(x)
    if velocity:
        for n, v in velocity.items():
            dx[n] = np.asarray(v, np.float32) * np.float32(h)
    T = O.ContactTwin(S, x, dx, flags=flags)
    _, _, pairs, cmeta = T.detect()
    out = dx.copy()
    for n in range(S.node_count):
        step = dx[n] + (np.asarray(delta.get(n, (0, 0, 0)), np.float32))
        out[n] = O.feasible(cmeta[n], step)
    return out


def test_bounds_keep_node_and_face_apart():
    S, x = _probe_scene()
    nx = 64
    f = 20 * nx + 20
    cs = [f, f + 1, f + nx, f + nx + 1]
    node = 50 * nx + 50
    face = {c: [((q % 2) - 0.5) * 0.00342, ((q // 2) - 0.5) * 0.00348, 0.0] for q, c in enumerate(cs)}
    above = dict(face)
    above[node] = [0.0, 0.0, 0.0015]
    gap = lambda out: 0.0015 + out[node, 2] - np.mean([out[c, 2] for c in cs])  # noqa: E731
    cases = {
        "a node driven through a still face": {node: (0, 0, -0.002)},
        "a face driven through a still node": {c: (0, 0, 0.003) for c in cs},
        "node and face driven through each other": {node: (0, 0, -0.002), **{c: (0, 0, 0.002) for c in cs}},
    }
    for name, delta in cases.items():
        g = gap(_apply(S, x, above, delta))
        assert g > 0.3e-3, (name, g)
    # 2.4 m/s through the face in one substep: stopped by the face its path crosses
    g = gap(_apply(S, x, above, {}, velocity={node: (0, 0, -2.4)}))
    assert g > 0.3e-3, g


def test_exclusion_window_links_own_cloth_but_not_a_torn_lip():
    S, x = _probe_scene()
    nx = 64
    f = 20 * nx + 20
    cs = [f, f + 1, f + nx, f + nx + 1]
    face = {c: [((q % 2) - 0.5) * 0.00342, ((q // 2) - 0.5) * 0.00348, 0.0] for q, c in enumerate(cs)}
    lip, link = f + 3 * nx, f + 2 * nx
    torn = dict(face)
    torn[lip] = [0.0, 0.0, 0.0015]
    torn[link] = [0.01, 0.0, 0.0]
    delta = {lip: (0, 0, -0.002)}
    linked = _apply(S, x, torn, delta)
    assert 0.0015 + linked[lip, 2] < 0      # its own linked cloth: passes
    flags = S.seg_flags.copy()
    flags[2 * (f + nx) + 1] |= 2            # cut the column between the lip and the face
    cut = _apply(S, x, torn, delta, flags=flags)
    assert 0.0015 + cut[lip, 2] > 0.3e-3    # a lip across a tear: held


def test_reciprocal_bound_mirrors_the_bounding_node():
    S, x, dx = _fold_state()
    T = O.ContactTwin(S, x, dx)
    _, _, pairs, cmeta = T.detect()
    own, rec = cmeta[:, 0], cmeta[:, 1]
    bounded = np.nonzero(own[:, 3] < 1e8)[0]
    assert len(bounded) > 0
    # every reciprocal direction is the negated normal of some bounded node, with its share to 10 um
    normals = {tuple(-own[n, :3]): own[n, 3] for n in bounded}
    for n in np.nonzero(rec[:, 3] < 1e8)[0]:
        share = normals.get(tuple(rec[n, :3]))
        assert share is not None
        assert share - 1e-5 <= rec[n, 3] <= share + 1e-9


def test_floor_bound_z():
    half, bulge, k = 0.35e-3, 0.3e-3, 1300.0
    m, g, fz = 4e-7, 9.81, -0.4e-3
    delta = half + bulge
    sW = np.sqrt(m * g * delta / k)
    # far above: unchanged; a step through the floor stops within the trust region and above half
    assert O.floor_bound_z(0.01, 0.0, -0.001, half, m, g, fz) == pytest.approx(-0.001)
    z = O.floor_bound_z(0.0, 0.0, -0.01, half, m, g, fz)
    s = max(delta - (0.0 - fz), 0.0)
    lowest = min(max(fz + delta - max(sW, 2 * s), fz + half), 0.0)
    assert z == pytest.approx(lowest, abs=1e-9)
    assert z >= fz + half - 1e-9

"""Pins the FP32 restatement of the reference's cross-layer contact (contact.wgsl, restated in
oracle/contact_oracle.c).

The reference has no CPU counterpart of contact.wgsl, so each piece is checked against what the
mathematics fixes: the hash holds every node once, in the slot of its cell; the K nearest nodes of
the other plies and the bound's nearest node against a brute-force search over every node (exact
wherever the nearest nodes lie inside the 27 cells the hash visits, i.e. within one 1 mm cell);
the penalty as the derivative of its quadratic energy and the stiffness as its Gauss-Newton
Hessian; Coulomb friction above the regularisation speed and viscous below it; and, driven through
the FP64 oracle, an upper ply pressed onto a lower one rests on it instead of passing through.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from bd.configs import advance_grips, config_plies, plies_scene

K = 4
NONE = 0xFFFFFFFF
CELL = np.float32(1e-3)
REACH = 0.25e-3


def pressed_state(S, height=0.45e-3, noise=2e-5, seed=5):
    """Ply 1 lowered to `height` above ply 0 (inside the 0.7 mm contact distance), small random steps."""
    rng = np.random.default_rng(seed)
    x = S.position.astype(np.float64).copy()
    off = S.layers[1].node_offset
    x[off:, 2] = height
    dx = rng.standard_normal(x.shape) * noise
    return x.astype(np.float32), dx.astype(np.float32)


def layer_index(S):
    lay = np.zeros(S.node_count, dtype=np.int64)
    for L in S.layers:
        lay[L.node_offset:L.node_offset + L.node_count] = L.index
    return lay


def test_hash_holds_every_node_once_in_its_cell():
    S = plies_scene(plies=3)
    x, dx = pressed_state(S)
    T = O.PlyTwin(S, x, dx)
    cells, items, pairs, cmeta = T.detect()
    T2 = T.TABLE
    starts, counts = cells[:T2], cells[T2:]
    assert counts.sum() == S.node_count
    assert np.array_equal(np.sort(items), np.arange(S.node_count))
    # starts are the exclusive scan of the counts
    assert np.array_equal(starts[1:], np.cumsum(counts, dtype=np.uint64)[:-1].astype(np.uint32))
    p = x + dx
    lay = layer_index(S)
    lib = O.lib()
    for n in range(0, S.node_count, 7):
        c = np.floor(p[n] / CELL).astype(np.int64)
        h = lib.oc_ply_hash_cell(int(c[0]), int(c[1]), int(c[2]), int(lay[n]))
        assert n in items[starts[h]:starts[h] + counts[h]]


def brute(S, p, n, lay, dist3):
    """Distances (f64 of the f32 positions) from node n to every node of the other layers, with
    their contact distances."""
    other = np.nonzero(lay != lay[n])[0]
    d = np.linalg.norm(p[other].astype(np.float64) - p[n].astype(np.float64), axis=1)
    s = lay[n] + lay[other]
    r = np.where(s == 1, dist3[0], np.where(s == 2, dist3[1], dist3[2])).astype(np.float64)
    return other, d, r


@pytest.mark.parametrize("plies", [2, 3])
def test_nearest_foreign_nodes_and_bound_against_brute_force(plies):
    S = plies_scene(plies=plies)
    x, dx = pressed_state(S)
    T = O.PlyTwin(S, x, dx)
    _, _, pairs, cmeta = T.detect()
    p = (x + dx).astype(np.float32)
    lay = layer_index(S)
    checked_k = checked_b = dup = 0
    for n in range(S.node_count):
        other, d, r = brute(S, p, n, lay, T.dist3)
        e = d - r
        order = np.lexsort((other, e))
        cand = order[d[order] < r[order] + REACH]
        got = [m for m in pairs[n] if m != NONE]
        # every listed candidate is a node of another layer within reach
        assert all(lay[m] != lay[n] for m in got)
        for m in got:
            assert np.linalg.norm(p[m].astype(np.float64) - p[n]) < T.dist3.max() + REACH + 1e-7
        # a slot shared by two of the 27 neighbour cells is visited twice (as in the reference), so a
        # node can fill two places: the distinct candidates are then the nearest len(set) ones
        uniq = sorted(set(int(m) for m in got))
        dup += len(uniq) < len(got)
        top = cand[:len(uniq) if len(uniq) < len(got) else K]
        # complete wherever the K nearest lie within one cell (always inside the 27 cells visited)
        if len(top) and d[top].max() < 0.999e-3 and (len(cand) <= len(top) or e[cand[len(top)]] - e[cand[len(top) - 1]] > 1e-7):
            assert uniq == sorted(other[top].tolist()), n
            checked_k += 1
        # the bound: share of the clearance to the nearest foreign node, along the direction from it
        j = int(np.argmin(e))
        if d[j] < 0.999e-3 and (len(e) < 2 or np.partition(e, 1)[1] - e[j] > 1e-7):
            share = 0.45 * (max(0.0, e[j]) + 0.5 * T.dist3[0])
            assert cmeta[n, 0, 3] == pytest.approx(share, rel=2e-5, abs=1e-9)
            u = (p[n].astype(np.float64) - p[other[j]]) / d[j]
            assert np.allclose(cmeta[n, 0, :3], u, atol=2e-4)
            checked_b += 1
        assert tuple(cmeta[n, 1]) == (0.0, 0.0, 0.0, np.float32(1e9))
    assert checked_k > S.node_count // 2 and checked_b > S.node_count // 2
    assert dup < S.node_count // 10


def test_far_plies_have_no_candidates_and_no_bound():
    S = plies_scene(plies=2)
    x = S.position.astype(np.float32).copy()
    x[S.layers[1].node_offset:, 2] = 0.5  # half a metre above: nothing in the 27 cells
    T = O.PlyTwin(S, x, np.zeros_like(x))
    _, _, pairs, cmeta = T.detect()
    assert (pairs == NONE).all()
    assert (cmeta[:, 0, 3] == np.float32(1e9)).all() and (cmeta[:, 0, :3] == 0).all()
    ext = T.terms(pairs, np.float32(1 / 120))
    assert (ext == 0).all()


def ply_energy(p, n, pairs_n, q, r):
    """sum over the node's active pairs of 1/2 k (r - |p - q|)^2 (the quadratic penalty)."""
    E = 0.0
    for m in pairs_n:
        if m == NONE:
            continue
        d = np.linalg.norm(p - q[m])
        if 1e-9 < d < r:
            E += 0.5 * 3000.0 * (r - d) ** 2
    return E


def test_penalty_is_the_gradient_of_its_energy_with_gauss_newton_stiffness():
    S = plies_scene(plies=2)
    x, _ = pressed_state(S)
    dx = np.zeros_like(x)
    T = O.PlyTwin(S, x, dx)
    _, _, pairs, _ = T.detect()
    h = np.float32(1 / 120)
    ext = T.terms(pairs, h)
    q = x.astype(np.float64)
    r = float(T.dist3[0])
    eps_u = 0.002 * float(h)
    active = 0
    for n in range(S.node_count):
        if pairs[n, 0] == NONE:
            continue
        g = np.zeros(3)
        for a in range(3):
            e = np.zeros(3)
            e[a] = 1e-9
            g[a] = (ply_energy(q[n] + e, n, pairs[n], q, r) - ply_energy(q[n] - e, n, pairs[n], q, r)) / 2e-9
        lam_sum = 0.0
        tr = 0.0
        for m in pairs[n]:
            if m == NONE:
                continue
            d = np.linalg.norm(q[n] - q[m])
            if 1e-9 < d < r:
                lam = 3000.0 * (r - d)
                lam_sum += lam
                # Gauss-Newton normal stiffness plus the viscous friction stiffness on both tangents
                tr += 3000.0 + 2 * 0.45 * lam / eps_u
        if lam_sum == 0:
            continue
        active += 1
        scale = max(np.abs(g).max(), 3000.0 * 1e-6)
        assert np.allclose(ext[n, :3], g, atol=2e-4 * scale), n
        Ht = ext[n, 3] + ext[n, 6] + ext[n, 8]
        assert Ht == pytest.approx(tr, rel=1e-4), n
    assert active > 300


def two_node_scene(lift, tangential):
    """Two plies of 4 x 4 nodes spread 3 mm apart (one pair within reach): node 5 of ply 1 sits
    `lift` above node 5 of ply 0, and steps `tangential` sideways."""
    S = plies_scene(nx=8, ny=8, plies=2)  # 64 + 16 nodes; positions are replaced below
    N = S.node_count
    x = np.zeros((N, 3), np.float32)
    for L in S.layers:
        for j in range(L.ny):
            for i in range(L.nx):
                n = L.node_offset + j * L.nx + i
                x[n] = (3e-3 * i, 3e-3 * j, 0.0 if L.index == 0 else 5e-3)
    n0 = 1 * S.layers[0].nx + 1
    n1 = S.layers[1].node_offset + 1 * S.layers[1].nx + 1
    x[n1] = x[n0] + np.array([0.0, 0.0, lift], np.float32)
    dx = np.zeros_like(x)
    dx[n1, 0] = tangential
    return S, x, dx, n0, n1


@pytest.mark.parametrize("tangential", [1e-3, 1e-6])
def test_friction_is_coulomb_above_the_regularisation_speed(tangential):
    h = np.float32(1 / 120)
    S, x, dx, n0, n1 = two_node_scene(0.4e-3, tangential)
    x = x - dx  # the pair's separation is measured at x + dx
    T = O.PlyTwin(S, x, dx)
    _, _, pairs, cmeta = T.detect()
    assert list(pairs[n1]) == [n0, NONE, NONE, NONE]
    assert list(pairs[n0]) == [n1, NONE, NONE, NONE]
    ext = T.terms(pairs, h)
    lam = 3000.0 * (0.7e-3 - 0.4e-3)
    # normal: the penalty pushes ply 1 up (gradient -lam along +z)
    assert ext[n1, 2] == pytest.approx(-lam, rel=1e-4)
    eps_u = 0.002 * float(h)
    ft = ext[n1, 0]
    if tangential > eps_u:
        assert ft == pytest.approx(0.45 * lam, rel=1e-4)  # Coulomb: mu lam along the slip
    else:
        assert ft == pytest.approx(0.45 * lam * tangential / eps_u, rel=1e-3)  # viscous
    # the partner sees the opposite relative step
    assert ext[n0, 0] == pytest.approx(-ft, rel=1e-4)
    assert ext[n0, 2] == pytest.approx(lam, rel=1e-4)


def oracle_ply_substep(cpu, S, params, contact=True):
    """One substep of the FP64 oracle with the FP32 restatement's detection and terms (the order of
    encodeSubstep: predict, detect, projectStart, per iteration terms then the Newton step)."""
    cpu.predict()
    x32 = cpu.x.reshape(-1, 3).astype(np.float32)
    if contact:
        T = O.PlyTwin(S, x32, cpu.dx.reshape(-1, 3).astype(np.float32))
        _, _, pairs, cmeta = T.detect()
        cpu.set_cmeta(cmeta.astype(np.float64))
        cpu.project_start()
    for it in range(params.newton_iterations):
        if contact:
            T = O.PlyTwin(S, x32, cpu.dx.reshape(-1, 3).astype(np.float32))
            cpu.set_ext(T.terms(pairs, np.float32(params.h)).astype(np.float64))
        cpu.newton_step(it == 0)
    cpu.finalize()


def ply_gap(S, x):
    """Per upper-ply node over the lower ply's interior: its height above the nearest lower-ply node."""
    L0, L1 = S.layers
    lower = x[:L0.node_count]
    upper = x[L1.node_offset:]
    inner = (np.abs(upper[:, 0]) < 6e-3) & (np.abs(upper[:, 1]) < 7e-3)
    g = []
    for p in upper[inner]:
        k = np.argmin(np.linalg.norm(lower[:, :2] - p[:2], axis=1))
        g.append(p[2] - lower[k, 2])
    return np.array(g)


def test_upper_ply_pressed_onto_the_lower_one():
    """A centre grip presses the upper ply 1.5 mm down (its start is 0.75 mm above the lower ply):
    with contact it pushes the lower ply ahead of it and stays near the 0.7 mm contact distance;
    without contact it passes through."""
    cfg = config_plies()
    S, params = cfg["scene"], cfg["params"]
    runs = {}
    for contact in (True, False):
        grips = [dict(g) for g in cfg["grips"]]
        cpu = O.Oracle(S, params)
        for _ in range(15):
            advance_grips(grips, cfg["velocity"])
            cpu.set_grips(grips)
            oracle_ply_substep(cpu, S, params, contact)
        runs[contact] = ply_gap(S, cpu.x.reshape(-1, 3))
    g, gf = runs[True], runs[False]
    assert np.isfinite(g).all() and len(g) > 200
    assert g.min() > 0.3e-3 and np.median(g) < 0.9e-3, (g.min(), np.median(g))
    assert gf.min() < -0.3e-3, gf.min()

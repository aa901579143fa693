"""Detached pieces on the host and in the FP64 oracle (no GPU).

* The topology search (bd_topology_search, restating topology.ts:142-238) against scipy's connected
  components: freed = cloth nodes in no component (through active spans between cloth nodes) that
  holds a mount or a cloth anchor; pieces = components of the detached nodes through active spans,
  numbered by their lowest node, each contiguous and starting at that node.
* The oracle's curved line search (energy.wgsl:65-73): with zero rotations it equals the straight
  one, and along an exact rigid rotation of a free sheet every internal term is unchanged, so the
  energy change is the inertial one alone, while the straight (tangent) step stretches the spans.
"""
import ctypes as C

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

import oracle as O
from bd import capi
from bd.configs import sheet_scene
from bd.material import linear_sheet_material


def search(layers, bits, mounted, anchors=()):
    lib = capi.load_library()
    N = len(bits)
    nx = np.array([l[0] for l in layers], np.int32)
    ny = np.array([l[1] for l in layers], np.int32)
    off = np.array([l[2] for l in layers], np.int32)
    a = np.array(list(anchors) or [0], np.int32)
    freed = np.zeros(N, np.uint8)
    of = np.zeros(N, np.uint32)
    start = np.zeros(N + 1, np.uint32)
    nodes = np.zeros(N + 1, np.uint32)
    count = C.c_int32()
    i32, u8, u32 = C.POINTER(C.c_int32), C.POINTER(C.c_uint8), C.POINTER(C.c_uint32)
    b = np.ascontiguousarray(bits, np.uint8)
    m = np.ascontiguousarray(mounted, np.uint8)
    rc = lib.bd_topology_search(N, len(layers), nx.ctypes.data_as(i32), ny.ctypes.data_as(i32), off.ctypes.data_as(i32),
                                b.ctypes.data_as(u8), m.ctypes.data_as(u8), a.ctypes.data_as(i32), len(anchors),
                                freed.ctypes.data_as(u8), of.ctypes.data_as(u32), start.ctypes.data_as(u32),
                                nodes.ctypes.data_as(u32), C.byref(count))
    assert rc >= 0
    return rc, freed, of, start[:count.value + 1], nodes


def span_edges(layers, bits, keep):
    """Active spans between nodes that both satisfy `keep`."""
    e = []
    for (nx, ny, off) in layers:
        for j in range(ny):
            for i in range(nx):
                n = off + j * nx + i
                if i + 1 < nx and bits[n] & 1 and keep[n] and keep[n + 1]:
                    e.append((n, n + 1))
                if j + 1 < ny and bits[n] & 2 and keep[n] and keep[n + nx]:
                    e.append((n, n + nx))
    return e


def components(N, edges):
    if not edges:
        return np.arange(N)
    a = np.array(edges)
    g = coo_matrix((np.ones(len(a)), (a[:, 0], a[:, 1])), shape=(N, N))
    return connected_components(g, directed=False)[1]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_topology_search_against_connected_components(seed):
    rng = np.random.default_rng(seed)
    layers = [(30, 20, 0), (12, 9, 600)]
    N = 600 + 108
    bits = np.zeros(N, np.uint8)
    for (nx, ny, off) in layers:
        for j in range(ny):
            for i in range(nx):
                n = off + j * nx + i
                if i + 1 < nx and rng.random() > 0.22:
                    bits[n] |= 1
                if j + 1 < ny and rng.random() > 0.22:
                    bits[n] |= 2
    # some nodes already detached, some retired
    pre = rng.random(N) < 0.05
    bits[pre] |= 4
    ret = pre & (rng.random(N) < 0.4)
    bits[ret] |= 8
    mounted = np.zeros(N, np.uint8)
    mounted[0:30] = 1  # the first row of layer 0
    anchors = [int(v) for v in rng.integers(0, N, 6)]
    nfree, freed, of, start, nodes = search(layers, bits, mounted, anchors)
    cloth = (bits & 4) == 0
    comp = components(N, span_edges(layers, bits, cloth))
    held = set(comp[(mounted == 1) & cloth]) | {comp[a] for a in anchors if cloth[a]}
    want_freed = cloth & ~np.isin(comp, list(held))
    assert np.array_equal(freed.astype(bool), want_freed) and nfree == want_freed.sum()
    detached = ((bits & 4) != 0) & ((bits & 8) == 0) | want_freed
    comp2 = components(N, span_edges(layers, bits, detached))
    labels = sorted({comp2[n] for n in np.nonzero(detached)[0]}, key=lambda c: np.nonzero(comp2 == c)[0][0])
    assert len(start) - 1 == len(labels)
    for p, c in enumerate(labels):
        members = np.nonzero((comp2 == c) & detached)[0]
        got = nodes[start[p]:start[p + 1]]
        assert got[0] == members[0] and sorted(got.tolist()) == members.tolist()
        assert (of[members] == p).all()
    assert (of[~detached] == 0xFFFFFFFF).all()


def test_topology_search_cut_out_block():
    """A block cut out of a mounted sheet is freed as one piece; a grip anchor inside it keeps it."""
    nx = ny = 16
    bits = np.full(nx * ny, 3, np.uint8)
    for j in range(ny):
        bits[j * nx + nx - 1] &= np.uint8(0xFE)
    bits[(ny - 1) * nx:] &= np.uint8(0xFD)
    inside = lambda i, j: 5 <= i < 10 and 4 <= j < 9
    for j in range(ny):
        for i in range(nx):
            n = j * nx + i
            if inside(i, j) != inside(i + 1, j):
                bits[n] &= np.uint8(0xFE)
            if inside(i, j) != inside(i, j + 1):
                bits[n] &= np.uint8(0xFD)
    mounted = np.zeros(nx * ny, np.uint8)
    mounted[(ny - 1) * nx:] = 1
    nfree, freed, of, start, nodes = search([(nx, ny, 0)], bits, mounted)
    block = [j * nx + i for j in range(4, 9) for i in range(5, 10)]
    assert nfree == 25 and sorted(np.nonzero(freed)[0].tolist()) == block
    assert len(start) == 2 and sorted(nodes[:25].tolist()) == block
    nfree, freed, *_ = search([(nx, ny, 0)], bits, mounted, anchors=[6 * nx + 7])
    assert nfree == 0 and not freed.any()


def free_sheet(n=10, pitch=3e-3):
    mat = dict(linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 8e3))
    mat["airDrag"] = 0.0
    return sheet_scene(n, pitch, mat, perturbation=0.0, pin_top=False), mat


def test_curved_line_search_without_rotation_is_the_straight_one():
    S, _ = free_sheet()
    params = capi.make_params(history=False)
    cpu = O.Oracle(S, params)
    cpu.predict()
    rng = np.random.default_rng(3)
    d = rng.standard_normal(3 * S.node_count) * 1e-4
    e0, m0 = cpu.energy_change(d)
    curve = np.zeros((S.node_count, 7))
    curve[:, 3] = 1.0
    curve[:, 4:] = rng.standard_normal((S.node_count, 3)) * 1e-2
    cpu.set_curve(curve)
    e1, m1 = cpu.energy_change(d)
    cpu.set_curve(None)
    assert np.allclose(e1, e0, rtol=1e-9, atol=1e-18) and np.allclose(m1, m0, rtol=1e-9, atol=1e-18)


def test_curved_line_search_along_a_rigid_rotation_changes_only_inertia():
    S, _ = free_sheet()
    params = capi.make_params(history=False)
    cpu = O.Oracle(S, params)
    cpu.predict()
    N = S.node_count
    x = cpu.x.reshape(-1, 3)
    m = np.asarray(S.mass)
    c = (m[:, None] * x).sum(0) / m.sum()
    r = x - c
    w = np.array([0.3, -0.2, 0.9])
    d = np.cross(w, r).ravel()  # the rotation's tangent
    curve = np.zeros((N, 7))
    curve[:, :3] = w
    curve[:, 3] = 1.0
    curve[:, 4:] = r
    cpu.set_curve(curve)
    e, _ = cpu.energy_change(d)
    cpu.set_curve(None)
    es, _ = cpu.energy_change(d)
    h = params.h
    pred = cpu.pred.reshape(-1, 3)
    d0 = cpu.dx.reshape(-1, 3)
    want = []
    for a in (1.0, 0.5, 0.25, 0.125):
        th = a * w
        phi = np.linalg.norm(th)
        k = th / phi
        Rr = r * np.cos(phi) + np.cross(k, r) * np.sin(phi) + np.outer(r @ k, k) * (1 - np.cos(phi))
        s = Rr - r
        want.append(np.sum(m / h ** 2 * (np.einsum("ij,ij->i", s, d0 - pred) + 0.5 * np.einsum("ij,ij->i", s, s))))
    want = np.array(want)
    assert np.allclose(e, want, rtol=1e-7), (e, want)
    # the straight tangent step stretches every span: more energy than the curved one
    assert (es > e + 1e-3 * np.abs(e)).all()


def test_topology_search_ignores_span_bits_at_the_lattice_edge():
    """A +u bit on a layer's last column or a +v bit on its last row names no neighbour: the search
    gives the same result as without those bits (and reads nothing past the layer)."""
    layers = [(7, 5, 0), (4, 3, 35)]
    N = 35 + 12
    rng = np.random.default_rng(9)
    bits = (rng.random(N) > 0.3).astype(np.uint8) | ((rng.random(N) > 0.3).astype(np.uint8) << 1)
    clean = bits.copy()
    for (nx, ny, off) in layers:
        for j in range(ny):
            for i in range(nx):
                n = off + j * nx + i
                if i == nx - 1:
                    clean[n] &= ~np.uint8(1)
                if j == ny - 1:
                    clean[n] &= ~np.uint8(2)
    mounted = np.zeros(N, np.uint8)
    mounted[:7] = 1
    a = search(layers, bits, mounted, [40])
    b = search(layers, clean, mounted, [40])
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x), np.asarray(y))

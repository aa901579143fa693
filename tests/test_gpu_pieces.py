"""Detached pieces on the GPU (csrc/pieces.hip, the assembly's rigid blocks, the PCG's rigid start
and the curved line search) through the C-ABI.

Numerics, each against a closed form or a numpy restatement:
* the rigid blocks of a resting cut-out piece (drag off): the external block is m / h^2 I, the
  centroid the mass centroid, and E^-1 inverts E = sum Z^T K Z (+ the rotation blocks);
* the PCG start is the piece's exact rigid response x0 = Z E^-1 Z^T b, zero elsewhere;
* the spin start of a spinning piece moves its Newton start onto the rotation continued over the
  substep, dx = h v + 2 (1 - cos w h) k x (k x r);
* the curved line-search energies match the FP64 oracle's (tests/gpu/correctness.ts bar 1e-2), and
  the curve apply bends delta by rotationCurve(alpha w, r) / alpha.
Behaviour, as the reference's tests/gpu/detached.ts checks it:
* a comb and a serpentine cut out of the mounted sheet fall at more than half the free-fall speed
  after 0.2 s; without the correction the same cut-out hovers;
* a free 4 x 4 piece thrown with a spin keeps its shape (strain < 2 %), its spin (> 0.8) and falls;
* two detached nodes pulled apart along their span: strain < 1 % over 0.25 s;
* a detached piece that falls below 0.25 m is retired: frozen, its spans removed.
"""
import sys

import numpy as np
import pytest

import oracle as O
from bd.capi import BD_BUF_DELTA, BD_BUF_RHS, BD_BUF_SEGFLAGS, BD_BUF_DX, make_params
from bd.configs import sheet_scene
from bd.material import linear_sheet_material
from bd.scene import SEG_BROKEN, SEG_EXISTS

pytestmark = pytest.mark.gpu

PITCH = 3.4e-3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def material(drag=True):
    mat = dict(linear_sheet_material(PITCH, 2e4, 1.6, 1e-5, 0.35, 8e3))
    if not drag:
        mat["airDrag"] = 0.0
    return mat


def cut_flags(S, inside, shape=None):
    """Span flags with every span across the outline of `inside` (a predicate on (i, j)) broken, and
    the inner cuts of a comb or a serpentine (detached.ts:309-343)."""
    L = S.layers[0]
    nx, ny = L.nx, L.ny
    flags = S.seg_flags.copy()
    rows = sorted({j for j in range(ny) for i in range(nx) if inside(i, j)})
    cols = sorted({i for j in range(ny) for i in range(nx) if inside(i, j)})
    j0, i0 = rows[0], cols[0]
    for j in range(ny):
        for i in range(nx):
            n = j * nx + i
            if i + 1 < nx and inside(i, j) != inside(i + 1, j):
                flags[2 * n] |= SEG_BROKEN
            if j + 1 < ny and inside(i, j) != inside(i, j + 1):
                flags[2 * n + 1] |= SEG_BROKEN
            if not inside(i, j):
                continue
            if shape == "comb" and j > j0 and inside(i + 1, j):
                flags[2 * n] |= SEG_BROKEN
            if shape == "serpentine" and inside(i, j + 1) and i != (cols[-1] if (j - j0) % 2 == 0 else i0):
                flags[2 * n + 1] |= SEG_BROKEN
    return flags


def block(i0, j0, w, h):
    return lambda i, j: i0 <= i < i0 + w and j0 <= j < j0 + h


def piece_nodes(S, inside):
    nx = S.layers[0].nx
    return [j * nx + i for j in range(S.layers[0].ny) for i in range(nx) if inside(i, j)]


def rigid_E(x, m, K, Kr, nodes):
    """E = sum Z^T K Z + rotation blocks about the mass centroid (rigid_frame.wgsl:96-124)."""
    c = (m[nodes, None] * x[nodes]).sum(0) / m[nodes].sum()
    E = np.zeros((6, 6))
    for n in nodes:
        r = x[n] - c
        Z = np.zeros((3, 6))
        Z[:, :3] = np.eye(3)
        for i in range(3):
            Z[:, 3 + i] = np.cross(np.eye(3)[i], r)
        E += Z.T @ K[n] @ Z
        E[3:, 3:] += Kr[n]
    return c, E


def sym(a):
    return np.array([[a[0], a[1], a[2]], [a[1], a[3], a[4]], [a[2], a[4], a[5]]])


def test_rigid_frame_and_start_of_a_cut_out_piece():
    from bd.solver import GpuSolver
    S = sheet_scene(24, PITCH, material(drag=False), perturbation=0.0)
    inside = block(8, 6, 8, 7)
    params = make_params(history=False, pcg_iterations=0)
    gpu = GpuSolver(S, params)
    gpu.write(BD_BUF_SEGFLAGS, cut_flags(S, inside))
    say("gpu: topology of an 8 x 7 block cut out of the hanging sheet")
    freed = gpu.update_topology()
    nodes = piece_nodes(S, inside)
    st = gpu.state()
    assert freed == len(nodes) and (st[nodes] == 1).all() and st.sum() == len(nodes)
    gpu.encode_predict()
    gpu.encode_assemble(True)
    x = gpu.x + gpu.dx
    m = np.asarray(S.mass, np.float64)
    rk = gpu.rigidk()
    w = np.float32(np.float32(m[0]) / np.float32(params.h) ** 2)
    assert np.allclose(rk[nodes, 0], [w, 0, 0, w, 0, w], rtol=1e-6), "external block = inertia"
    K = np.array([sym(rk[n, 0]) for n in range(S.node_count)])
    Kr = np.array([sym(rk[n, 1]) for n in range(S.node_count)])
    c, E = rigid_E(x, m, K, Kr, nodes)
    P = gpu.pieces()
    assert P["valid"][0] == 1 and np.allclose(P["centroid"][0], c, atol=1e-9)
    st_ = np.trace(E[:3, :3]) / 3
    sr = np.trace(E[3:, 3:]) / 3
    Er = E + np.diag([0, 0, 0] + [1e-6 * max(sr, 1e-8 * st_)] * 3)
    err = np.abs(P["Einv"][0] @ Er - np.eye(6)).max()
    say(f"E^-1 E - I: {err:.2e}")
    assert err < 1e-6
    say("gpu: PCG start (no iterations): the piece's rigid response")
    gpu.encode_pcg()
    b = gpu.read(BD_BUF_RHS).reshape(-1, 3)
    r = x[nodes] - c
    g = np.concatenate([b[nodes].sum(0), np.cross(r, b[nodes]).sum(0)])
    q = P["Einv"][0] @ g
    x0 = q[:3] + np.cross(q[3:], r)
    d = gpu.read(BD_BUF_DELTA).reshape(-1, 3)
    P = gpu.pieces()
    assert np.allclose(P["t"][0], q[:3], rtol=1e-6, atol=1e-12) and np.allclose(P["w"][0], q[3:], rtol=1e-6, atol=1e-12)
    e = np.abs(d[nodes] - x0).max() / np.abs(x0).max()
    say(f"rigid start vs Z E^-1 Z^T b: {e:.2e}; gravity step {x0[:, 2].mean() * 1e3:.3f} mm")
    assert e < 1e-6
    others = np.setdiff1d(np.arange(S.node_count), nodes)
    assert (d[others] == 0).all()


def isolate(S, placed, spans, velocity=None):
    """Only `placed` nodes in the scene (every other node retired) and only `spans` (owner, family)
    intact (detached.ts:36-61): returns the state, flags, positions and velocities to load."""
    N = S.node_count
    x = S.position.copy()
    state = np.full(N, 2, np.uint8)
    for n, p in placed.items():
        x[n] = p
        state[n] = 0
    flags = S.seg_flags | SEG_BROKEN
    for n, f in spans:
        flags[2 * n + f] = SEG_EXISTS
    v = np.zeros((N, 3))
    for n, u in (velocity or {}).items():
        v[n] = u
    return state, flags.astype(np.uint16), x, v


def load(gpu, state, flags, x, v):
    gpu.set_state(x, v)
    gpu.write(BD_BUF_SEGFLAGS, flags)
    gpu.set_node_state(state)
    return gpu.update_topology()


def tumble_setup(S, omega=10.0):
    nx = S.layers[0].nx
    c0 = 8 * nx + 8
    placed, vel, spans = {}, {}, []
    for j in range(4):
        for i in range(4):
            n = c0 + j * nx + i
            p = np.array([(i - 1.5) * PITCH, (j - 1.5) * PITCH, 0.1])
            placed[n] = p
            vel[n] = np.array([-omega * p[1], omega * p[0], 0.0])
            if i < 3:
                spans.append((n, 0))
            if j < 3:
                spans.append((n, 1))
    return placed, vel, spans


def strain(S, x, spans):
    nx = S.layers[0].nx
    return max(abs(np.linalg.norm(x[n + (1 if f == 0 else nx)] - x[n]) / S.rest_length[2 * n + f] - 1) for n, f in spans)


def test_spin_start_follows_the_rotation():
    from bd.solver import GpuSolver
    S = sheet_scene(20, PITCH, material(), perturbation=0.0)
    params = make_params(history=False)
    gpu = GpuSolver(S, params)
    placed, vel, spans = tumble_setup(S)
    freed = load(gpu, *isolate(S, placed, spans, vel))
    assert freed == 16
    gpu.encode_predict()
    nodes = sorted(placed)
    x = np.array([placed[n] for n in nodes])
    v = np.array([vel[n] for n in nodes])
    c = x.mean(0)
    h = params.h
    phi = 10.0 * h
    k = np.array([0, 0, 1.0])
    r = x - c
    kkr = np.cross(k, np.cross(k, r))
    want = h * v + 2 * (1 - np.cos(phi)) * kkr
    got = gpu.dx[nodes]
    e = np.abs(got - want).max() / np.abs(want).max()
    say(f"spin start vs closed form: {e:.2e}")
    assert e < 1e-5


def test_curved_line_search_matches_the_oracle():
    from bd.solver import GpuSolver
    S = sheet_scene(20, PITCH, material(), perturbation=0.0)
    params = make_params(history=False)
    gpu = GpuSolver(S, params)
    placed, vel, spans = tumble_setup(S)
    state, flags, x, v = isolate(S, placed, spans, vel)
    load(gpu, state, flags, x, v)
    gpu.encode_predict()
    gpu.encode_assemble(True)
    gpu.encode_pcg()
    gpu.encode_line_search()
    cv = gpu.curve()
    nodes = sorted(placed)
    P = gpu.pieces()
    xi = gpu.x + gpu.dx
    assert (cv[nodes, 3] == 1).all() and (np.delete(cv[:, 3], nodes) == 0).all()
    assert np.allclose(cv[nodes, :3], P["w"][0], rtol=1e-6)
    assert np.allclose(cv[nodes, 4:], xi[nodes] - P["centroid"][0], atol=1e-7)
    eg, mg = gpu.energy()
    d = gpu.read(BD_BUF_DELTA)
    cpu = O.Oracle(S, params)
    cpu.x[:] = gpu.x.ravel()
    cpu.v[:] = gpu.v.ravel()
    cpu.predict()
    cpu.dx[:] = gpu.dx.ravel()
    cpu.seg_flags[:] = flags
    cpu.pinned[state == 2] = 1  # retired nodes are out of the solve
    cpu.set_curve(cv.astype(np.float64))
    ec, mc = cpu.energy_change(d)
    err = np.max(np.abs(eg - ec)) / np.max(np.abs(ec))
    say(f"curved line-search energies gpu {eg} cpu {ec} rel {err:.2e}")
    assert err < 1e-2
    assert np.max(np.abs(mg - mc) / np.abs(mc)) < 1e-2
    # the update then moves the piece's nodes along the chosen step's rotation
    alpha = gpu.line_alpha()
    d0 = gpu.read(BD_BUF_DELTA).reshape(-1, 3)
    gpu.encode_newton_apply(0)
    d1 = gpu.read(BD_BUF_DELTA).reshape(-1, 3)
    if alpha > 0:
        th = alpha * P["w"][0]
        phi = np.linalg.norm(th)
        k = th / phi
        r = cv[nodes, 4:].astype(np.float64)
        kr = np.cross(k, r)
        bend = ((np.sin(phi) - phi) * kr + (1 - np.cos(phi)) * np.cross(k, kr)) / alpha
        assert np.allclose(d1[nodes] - d0[nodes], bend, atol=1e-7 + 1e-4 * np.abs(bend).max())


def fall_speed(S, inside, shape, pieces=True, steps=24):
    from bd.solver import GpuSolver
    params = make_params(history=False)
    gpu = GpuSolver(S, params)
    gpu.write(BD_BUF_SEGFLAGS, cut_flags(S, inside, shape))
    if pieces:
        gpu.update_topology()
    for _ in range(steps):
        gpu.encode_substep()
    nodes = piece_nodes(S, inside)
    return -gpu.v[nodes, 2].mean(), steps * params.h, gpu.state()[nodes]


@pytest.mark.parametrize("shape", ["comb", "serpentine"])
def test_cut_out_piece_falls(shape):
    S = sheet_scene(48, PITCH, material(), perturbation=0.0)
    inside = block(20, 20, 8, 7)
    # 0.2 s: the piece then has fallen 0.2 m, and it leaves the scene (retired) at 0.25 m
    say(f"gpu: a {shape} cut out of the hanging sheet, 0.2 s")
    vz, t, st = fall_speed(S, inside, shape)
    vz0, _, _ = fall_speed(S, inside, shape, pieces=False)
    say(f"  fall {vz:.3f} m/s with the rigid-mode correction, {vz0:.3f} m/s without (free fall {9.81 * t:.3f})")
    assert (st >= 1).all()
    assert vz > 0.5 * 9.81 * t
    assert vz0 < vz


def test_spinning_piece_keeps_shape_and_spin():
    from bd.solver import GpuSolver
    S = sheet_scene(20, PITCH, material(), perturbation=0.0)
    params = make_params(history=False)
    gpu = GpuSolver(S, params)
    placed, vel, spans = tumble_setup(S)
    load(gpu, *isolate(S, placed, spans, vel))
    nodes = sorted(placed)
    worst = 0.0
    steps = 5 * 6
    for _ in range(6):
        for _ in range(5):
            gpu.encode_substep()
        worst = max(worst, strain(S, gpu.x, spans))
    x, v = gpu.x[nodes], gpu.v[nodes]
    com, vcom = x.mean(0), v.mean(0)
    r = x - com
    spin = np.sum(r[:, 0] * (v[:, 1] - vcom[1]) - r[:, 1] * (v[:, 0] - vcom[0])) / np.sum(r[:, 0] ** 2 + r[:, 1] ** 2)
    t = steps * params.h
    say(f"tumble: strain {worst * 100:.2f} %, spin {spin:.2f} of 10 rad/s, fall {-vcom[2]:.3f} m/s (free {9.81 * t:.3f})")
    assert (gpu.state()[nodes] == 1).all()
    assert worst < 0.02 and spin > 8.0 and -vcom[2] > 0.5 * 9.81 * t


def test_detached_span_resists_extension():
    from bd.solver import GpuSolver
    S = sheet_scene(20, PITCH, material(), perturbation=0.0)
    params = make_params(history=False, gravity=(0.0, 0.0, 0.0))
    gpu = GpuSolver(S, params)
    a = 8 * 20 + 8
    L = S.rest_length[2 * a]
    placed = {a: np.array([-L / 2, 0, 0.08]), a + 1: np.array([L / 2, 0, 0.08])}
    vel = {a: np.array([-0.1, 0, 0]), a + 1: np.array([0.1, 0, 0])}
    assert load(gpu, *isolate(S, placed, [(a, 0)], vel)) == 2
    worst = 0.0
    for _ in range(6):
        for _ in range(5):
            gpu.encode_substep()
        worst = max(worst, strain(S, gpu.x, [(a, 0)]))
    say(f"differential: largest strain {worst * 100:.3f} %")
    assert worst < 0.01 and gpu.state()[a] == 1
    assert gpu.read(BD_BUF_SEGFLAGS)[2 * a] == SEG_EXISTS


def test_piece_below_the_scene_is_retired():
    from bd.solver import GpuSolver
    S = sheet_scene(20, PITCH, material(), perturbation=0.0)
    params = make_params(history=False)
    gpu = GpuSolver(S, params)
    placed, vel, spans = tumble_setup(S, omega=0.0)
    for n in placed:
        placed[n][2] = -0.24
    load(gpu, *isolate(S, placed, spans, vel))
    nodes = sorted(placed)
    for _ in range(20):
        gpu.encode_substep()
    st = gpu.state()[nodes]
    flags = gpu.read(BD_BUF_SEGFLAGS)
    x1 = gpu.x[nodes].copy()
    gpu.encode_substep()
    say(f"retired {int((st == 2).sum())} of {len(nodes)} nodes, lowest z {x1[:, 2].min():.3f} m")
    assert (st == 2).all()
    assert all((flags[2 * n + f] & SEG_EXISTS) == 0 for n, f in spans)
    assert np.array_equal(gpu.x[nodes], x1) and (gpu.v[nodes] == 0).all()

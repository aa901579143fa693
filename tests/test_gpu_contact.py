"""Self-contact and floor on the GPU through the C-ABI against the FP32 restatement
(oracle/contact_oracle.c) and the FP64 oracle.

* detection (hash, scan, fill, row marks, K nearest, bounds, reciprocal bounds) and the terms are
  compared bit for bit: both evaluate the same FP32 expressions with every product rounded as
  written, and the scatter sums are fixed-point integers whose order does not matter;
* the bound behaviour of the reference's tests/gpu/interaction.ts:185-248, through bd_newton_apply;
* assembly with contact terms and active bounds against the FP64 oracle (the reference's fine
  operator and rhs bars, 5e-5 and 1e-3);
* a flap turned over onto the cloth, dropped onto the floor: it lands on the lower flap and stays
  above it (and falls into it without self-contact); the GPU trajectory follows the oracle driven
  by the FP32 restatement within the reference's dynamics bar.
"""
import sys

import numpy as np
import pytest

import oracle as O
from bd.capi import BD_BUF_DELTA, BD_BUF_DX, BD_BUF_RHS, make_params
from bd.configs import config_flaps, config_fold, sheet_scene
from bd.material import linear_sheet_material

pytestmark = pytest.mark.gpu
K = 4


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def f32_gnorm(g):
    g = np.asarray(g, dtype=np.float32)
    return np.float32(np.sqrt((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]))


def fold_state(push=2.5e-3, noise=1e-4, seed=3):
    cfg = config_fold()
    S = cfg["scene"]
    x = S.position.astype(np.float32)
    dx = np.zeros_like(x)
    dx[x[:, 2] > 1.5e-3, 2] = -push
    dx += (np.random.default_rng(seed).standard_normal(dx.shape) * noise).astype(np.float32)
    return cfg, S, x, dx


def sorted_records(cells, items, T2=1 << 17):
    """Per occupied cell, its records sorted by face id (fill order is free)."""
    starts, counts = cells[:T2], cells[T2:]
    out = {}
    for h in np.nonzero(counts)[0]:
        r = items[starts[h]:starts[h] + counts[h]]
        out[int(h)] = r[np.argsort(r[:, 7].view(np.uint32), kind="stable")].view(np.uint32)
    return out


def test_detection_and_terms_bit_exact():
    from bd.solver import GpuSolver
    cfg, S, x, dx = fold_state()
    params = cfg["params"]
    say("gpu: contact detect on the pushed fold")
    gpu = GpuSolver(S, params)
    gpu.set_state(x.astype(np.float64))
    gpu.write(BD_BUF_DX, dx.astype(np.float64))
    gpu.encode_contact_detect()
    T = O.ContactTwin(S, x, dx)
    cells, items, pairs, cmeta = T.detect()
    gp, gc, gcells = gpu.pairs(), gpu.cmeta(), gpu.cells()
    assert np.array_equal(gcells, cells), "cell starts / counts"
    rg, rc = sorted_records(gcells, gpu.items()), sorted_records(cells, items)
    assert rg.keys() == rc.keys()
    for h in rc:
        assert np.array_equal(rg[h], rc[h]), f"records of cell {h}"
    assert np.array_equal(gp, pairs), "candidate faces, keys, row marks"
    assert np.array_equal(gc.view(np.uint32), cmeta.view(np.uint32)), "bounds"
    n_pairs = int((pairs[:, :K] != 0xFFFFFFFF).sum())
    say(f"detect: {n_pairs} candidates, {(cmeta[:, 0, 3] < 1e8).sum()} bounded, "
        f"{(cmeta[:, 1, 3] < 1e8).sum()} reciprocal")
    assert n_pairs > 0
    say("gpu: contact terms")
    gpu.encode_contact_terms()
    cm = cmeta.copy()
    acc, ext = T.terms(pairs, cm, np.float32(params.h), f32_gnorm(params.gravity))
    ge = gpu.ext()
    assert np.array_equal(ge.view(np.uint32), ext.view(np.uint32)), "ext"
    assert np.array_equal(gpu.cmeta().view(np.uint32), cm.view(np.uint32)), "trust-region bounds"
    assert (np.abs(ext[:, 2]) > 0).sum() > 100


def probe_solver(flags=None):
    from bd.solver import GpuSolver
    mat = linear_sheet_material(3.42e-3, 2e4, 1.6, 1e-5, 0.35, 8e3)
    S = sheet_scene(64, 3.42e-3, mat, perturbation=0.0, pin_top=False)
    x = S.position.astype(np.float32)
    x[:, 1] += 0.3
    params = make_params(self_contact=True, line_search=False, history=False)
    return S, x, GpuSolver(S, params)


def gpu_apply(gpu, x, placed, delta, velocity=None):
    x = x.copy()
    for n, p in placed.items():
        x[n] = p
    v = np.zeros_like(x, dtype=np.float64)
    for n, q in (velocity or {}).items():
        v[n] = q
    gpu.set_state(x.astype(np.float64), v)
    gpu.write(BD_BUF_DX, np.zeros(3 * len(x)))
    if velocity:
        gpu.encode_predict()
    d = np.zeros((len(x), 3))
    for n, q in delta.items():
        d[n] = q
    gpu.write(BD_BUF_DELTA, d)
    gpu.encode_line_search()      # line search off: alpha = 1
    gpu.encode_contact_detect()
    gpu.encode_newton_apply(0)
    return gpu.dx


def test_bounds_keep_node_and_face_apart():
    S, x, gpu = probe_solver()
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
        say(f"gpu: contact case: {name}")
        g = gap(gpu_apply(gpu, x, above, delta))
        say(f"  gap {g * 1e3:.3f} mm")
        assert g > 0.3e-3, name
    say("gpu: contact case: fast node")
    g = gap(gpu_apply(gpu, x, above, {}, velocity={node: (0, 0, -2.4)}))
    say(f"  gap {g * 1e3:.3f} mm")
    assert g > 0.3e-3
    # the lip across a tear inside the exclusion window
    lip, link = f + 3 * nx, f + 2 * nx
    torn = dict(face)
    torn[lip] = [0.0, 0.0, 0.0015]
    torn[link] = [0.01, 0.0, 0.0]
    say("gpu: contact case: linked lip")
    linked = gpu_apply(gpu, x, torn, {lip: (0, 0, -0.002)})
    assert 0.0015 + linked[lip, 2] < 0
    flags = S.seg_flags.copy()
    flags[2 * (f + nx) + 1] |= 2
    gpu.set_segments(seg_flags=flags)
    say("gpu: contact case: torn lip")
    cut = gpu_apply(gpu, x, torn, {lip: (0, 0, -0.002)})
    gpu.set_segments(seg_flags=S.seg_flags)
    assert 0.0015 + cut[lip, 2] > 0.3e-3


def rel_to_scale(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_assembly_with_contact_terms_and_floor():
    """Predicted fall of the fold's upper layer, made feasible (projectStart), then the contact terms
    and the assembly with active bounds (most upper nodes end at a bound with the inertia still
    driving them on) and the floor under the lower layer."""
    from bd.solver import GpuSolver
    cfg = config_fold()
    S, params = cfg["scene"], cfg["params"]
    x = S.position.astype(np.float32)
    v = np.zeros(x.shape)
    v[x[:, 2] > 1.5e-3, 2] = -0.3
    gpu = GpuSolver(S, params)
    gpu.set_floor(cfg["floor_z"])
    gpu.set_state(x.astype(np.float64), v)
    say("gpu: predict + detect + projectStart + terms + assemble on the falling fold")
    gpu.encode_predict()
    gpu.encode_contact_detect()
    gpu.encode_project_start()
    gpu.encode_contact_terms()
    gpu.encode_assemble(True)
    dxg, ext, cm = gpu.dx, gpu.ext(), gpu.cmeta()
    assert (np.abs(dxg[x[:, 2] > 1.5e-3, 2]) < 2.4e-3).all()  # the projected fall
    cpu = O.Oracle(S, params)
    cpu.x[:] = x.astype(np.float64).ravel()
    cpu.v[:] = v.ravel()
    cpu.predict()
    cpu.dx[:] = dxg.ravel()
    cpu.set_ext(ext.astype(np.float64))
    cpu.set_cmeta(cm.astype(np.float64))
    cpu.set_floor(cfg["floor_z"])
    cpu.assemble()
    Ag, Ac = gpu.stencil(0), cpu.stencil(0)
    e = rel_to_scale(Ag.reshape(len(Ag), -1), Ac.reshape(len(Ac), -1))
    say(f"fine operator with contact maxAbs/scale = {e:.2e}")
    assert e < 5e-5
    e = rel_to_scale(gpu.read(BD_BUF_RHS), cpu.rhs())
    say(f"rhs with contact maxAbs/scale = {e:.2e}")
    assert e < 1e-3


def oracle_contact_substep(cpu, S, params, floor_z):
    """One substep of the FP64 oracle with the FP32 restatement's detection and terms between its
    Newton iterations (the order of encodeSubstep: predict, detect, projectStart, per iteration
    terms then assemble/solve/line search/apply)."""
    N = S.node_count
    g = f32_gnorm(params.gravity)
    cpu.predict()
    x32 = cpu.x.reshape(-1, 3).astype(np.float32)
    T = O.ContactTwin(S, x32, cpu.dx.reshape(-1, 3).astype(np.float32))
    _, _, pairs, cmeta = T.detect()
    cpu.set_cmeta(cmeta.astype(np.float64))
    cpu.set_floor(floor_z)
    cpu.project_start()
    for it in range(params.newton_iterations):
        T = O.ContactTwin(S, x32, cpu.dx.reshape(-1, 3).astype(np.float32))
        _, ext = T.terms(pairs, cmeta, np.float32(params.h), g)
        cpu.set_ext(ext.astype(np.float64))
        cpu.set_cmeta(cmeta.astype(np.float64))
        cpu.newton_step(it == 0)
    cpu.finalize()
    assert cpu.x.shape[0] == 3 * N


def layer_gap(S, x):
    """Per upper-flap node (away from the edges): its height above the lower-flap cell it was
    turned over onto (mean of that cell's corners)."""
    nx, ny = S.layers[0].nx, S.layers[0].ny
    half = S.extra["flaps"]["half"]
    P = np.asarray(x).reshape(ny, nx, 3)
    g = []
    for j in range(half + 2, ny - 3):
        jl = ny - 1 - j
        for i in range(2, nx - 3):
            g.append(P[j, i, 2] - P[jl:jl + 2, i:i + 2, 2].mean())
    return np.array(g)


def test_flap_lands_on_the_cloth():
    from bd.solver import GpuSolver
    cfg = config_flaps()
    S, params, fz = cfg["scene"], cfg["params"], cfg["floor_z"]
    steps = 12
    gpu = GpuSolver(S, params)
    gpu.set_floor(fz)
    cpu = O.Oracle(S, params)
    x0 = cpu.x.copy().reshape(-1, 3)
    say("gpu: flap drop with self-contact and floor")
    parity = None
    for k in range(steps):
        gpu.encode_substep()
        oracle_contact_substep(cpu, S, params, fz)
        st = gpu.stats()
        say(f"  substep {k}: newton {st['newton_iterations']} pcg {st['pcg_iterations'][:st['newton_iterations']]}")
        if k == 3:
            # the fall and the landing (4 substeps): the trajectory against the oracle. Later substeps
            # are the penalty contacts' settling bounce, where the PCG iteration cap is reached and
            # the two solvers' unconverged iterates part; those are held to the gap checks below.
            xg4, xc4 = gpu.x, cpu.x.reshape(-1, 3).copy()
            parity = (np.max(np.linalg.norm(xg4 - xc4, axis=1)), np.max(np.linalg.norm(xc4 - x0, axis=1)))
    xg = gpu.x
    xc = cpu.x.reshape(-1, 3)
    half = 0.5 * S.layers[0].material["thickness"]
    assert np.isfinite(xg).all()
    assert xg[:, 2].min() >= fz + half - 1e-5, "floor"
    gaps = layer_gap(S, xg)
    say(f"layer gap min {gaps.min() * 1e3:.3f} mm median {np.median(gaps) * 1e3:.3f} mm ({len(gaps)} nodes)")
    assert len(gaps) > 100 and gaps.min() > 0.3e-3
    # the upper flap actually fell from 2 mm onto the lower one (held near the 1 mm contact distance)
    assert np.median(gaps) < 1.5e-3
    err, motion = parity
    say(f"gpu vs oracle after 4 substeps: max |dx| {err * 1e3:.4f} mm of motion {motion * 1e3:.3f} mm")
    assert err < 0.25 * motion
    err12 = np.max(np.linalg.norm(xg - xc, axis=1))
    say(f"gpu vs oracle after {steps} substeps: max |dx| {err12 * 1e3:.4f} mm")
    assert np.isfinite(xc).all() and err12 < 1e-3


def test_flap_passes_through_without_self_contact():
    from bd.solver import GpuSolver
    cfg = config_flaps()
    S, fz = cfg["scene"], cfg["floor_z"]
    params = make_params(h=1 / 120, newton_iterations=3, pcg_rel_tol=1e-5, pcg_max_iterations=100, history=False,
                         self_contact=False)
    gpu = GpuSolver(S, params)
    gpu.set_floor(fz)
    say("gpu: flap drop without self-contact")
    for _ in range(12):
        gpu.encode_substep()
    gaps = layer_gap(S, gpu.x)
    say(f"layer gap without self-contact: min {gaps.min() * 1e3:.3f} mm")
    assert gaps.min() < 0.3e-3


def test_fresh_solvers_detect_what_the_scene_holds():
    """Three fresh solvers of the default worn coupon (self-contact on): contact detection on the
    initial state matches the FP32 restatement bit for bit each time: the row marks of the
    coupon's broken spans, the candidate lists and the bounds. The buffers a solver uploads at
    creation (lattice coordinates, span flags, positions) must hold the scene before its first
    kernel runs: their allocations are cleared on the null stream, which the solver's non-blocking
    stream does not wait for, so each clear has to finish before the upload is queued."""
    from bd.scene import compile_scene, default_scene_specs
    from bd.solver import GpuSolver
    S = compile_scene(default_scene_specs())
    x = S.position.astype(np.float32)
    T = O.ContactTwin(S, x, np.zeros_like(x))
    _, _, pairs, cmeta = T.detect()
    marks = int(pairs[:, K + 1].sum())
    say(f"restatement: {marks} row marks, {(cmeta[:, 0, 3] < 1e8).sum()} bounded, "
        f"{(cmeta[:, 1, 3] < 1e8).sum()} reciprocal")
    assert marks > 0
    for rep in range(3):
        gpu = GpuSolver(S, make_params(self_contact=True))
        gpu.encode_contact_detect()
        gp, gc = gpu.pairs(), gpu.cmeta()
        say(f"solver {rep}: {int(gp[:, K + 1].sum())} row marks")
        assert np.array_equal(gp, pairs), f"solver {rep}: candidate faces, keys, row marks"
        assert np.array_equal(gc.view(np.uint32), cmeta.view(np.uint32)), f"solver {rep}: bounds"
        del gpu

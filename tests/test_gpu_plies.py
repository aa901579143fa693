"""Cross-layer contact of stacked plies on the GPU (csrc/plies.hip) through the C-ABI.

* Detection and terms on three pressed plies are bit-identical to the FP32 restatement of
  contact.wgsl (oracle/contact_oracle.c): cell starts and counts, each cell's node set, the K
  candidates, the bounds and ext.
* The assembly with these terms and bounds matches the FP64 oracle at the reference's bars
  (tests/gpu/correctness.ts: blocks 5e-5, rhs 1e-3).
* An upper ply pressed onto the lower one (both mounted by their collars) follows the oracle
  (driven by the restatement's contact) over 4 substeps to 0.25 of the motion (the reference's
  dynamics bar) and rests on the lower ply after 10, near the 0.7 mm contact distance.
"""
import sys

import numpy as np
import pytest

import oracle as O
from bd.capi import BD_BUF_DELTA, BD_BUF_DX, BD_BUF_LEVEL_B, BD_BUF_LEVEL_X, BD_BUF_RHS
from bd.configs import advance_grips, config_plies, plies_scene

pytestmark = pytest.mark.gpu

K = 4
T2 = 1 << 20


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pressed(S, step=0.45e-3, noise=2e-5, seed=9):
    """Ply l lowered to l x `step` (inside the 0.7 mm contact distance of its neighbours), small
    random steps."""
    x = S.position.astype(np.float64).copy()
    for L in S.layers[1:]:
        x[L.node_offset:L.node_offset + L.node_count, 2] = L.index * step
    dx = np.random.default_rng(seed).standard_normal(x.shape) * noise
    return x.astype(np.float32), dx.astype(np.float32)


def cell_sets(cells, items):
    starts, counts = cells[:T2], cells[T2:]
    return {int(h): sorted(items[starts[h]:starts[h] + counts[h]].tolist()) for h in np.nonzero(counts)[0]}


def test_ply_detection_and_terms_bit_exact():
    from bd.solver import GpuSolver
    cfg = config_plies(plies=3)
    S, params = cfg["scene"], cfg["params"]
    x, dx = pressed(S)
    say(f"gpu: cross-layer detect on 3 pressed plies ({S.node_count} nodes)")
    gpu = GpuSolver(S, params)
    assert gpu.contact_mode == "plies"
    gpu.set_state(x.astype(np.float64))
    gpu.write(BD_BUF_DX, dx.astype(np.float64))
    gpu.encode_contact_detect()
    T = O.PlyTwin(S, x, dx)
    cells, items, pairs, cmeta = T.detect()
    gcells = gpu.cells()
    assert np.array_equal(gcells, cells), "cell starts / counts"
    assert cell_sets(gcells, gpu.items()) == cell_sets(cells, items), "nodes per cell"
    gp = gpu.pairs()
    assert np.array_equal(gp, pairs), "K nearest foreign nodes"
    assert np.array_equal(gpu.cmeta().view(np.uint32), cmeta.view(np.uint32)), "bounds"
    n_pairs = int((pairs != 0xFFFFFFFF).sum())
    say(f"detect: {n_pairs} candidates, {(cmeta[:, 0, 3] < 1e8).sum()} bounded")
    assert n_pairs > 2 * S.node_count
    say("gpu: cross-layer terms")
    gpu.encode_contact_terms()
    ext = T.terms(pairs, np.float32(params.h))
    assert np.array_equal(gpu.ext().view(np.uint32), ext.view(np.uint32)), "ext"
    assert (ext[:, 2] != 0).sum() > S.node_count // 2


def rel_to_scale(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_assembly_with_ply_contact():
    """Predicted fall of two pressed plies made feasible (projectStart), then the contact terms and
    the assembly with active bounds. Both plies are stretched by 0.2 % in the plane so that no span
    sits at the axial law's kink at rest length, where FP32 and FP64 may take different branches
    (the rows the reference's own check excludes, tests/gpu/correctness.ts:55-68)."""
    from bd.solver import GpuSolver
    cfg = config_plies()
    S, params = cfg["scene"], cfg["params"]
    x, _ = pressed(S, step=0.6e-3)
    x[:, :2] *= np.float32(1.002)
    v = np.zeros(x.shape)
    v[S.layers[1].node_offset:, 2] = -0.1
    gpu = GpuSolver(S, params)
    gpu.set_state(x.astype(np.float64), v)
    say("gpu: predict + detect + projectStart + terms + assemble on two plies")
    gpu.encode_predict()
    gpu.encode_contact_detect()
    gpu.encode_project_start()
    gpu.encode_contact_terms()
    gpu.encode_assemble(True)
    dxg, ext, cm = gpu.dx, gpu.ext(), gpu.cmeta()
    assert (np.abs(ext[:, :3]).max(axis=1) > 0).sum() > 100
    cpu = O.Oracle(S, params)
    cpu.x[:] = x.astype(np.float64).ravel()
    cpu.v[:] = v.ravel()
    cpu.predict()
    cpu.dx[:] = dxg.ravel()
    cpu.set_ext(ext.astype(np.float64))
    cpu.set_cmeta(cm.astype(np.float64))
    cpu.assemble()
    Ag, Ac = gpu.stencil(0), cpu.stencil(0)
    e = rel_to_scale(Ag.reshape(len(Ag), -1), Ac.reshape(len(Ac), -1))
    say(f"fine operator with ply contact maxAbs/scale = {e:.2e}")
    assert e < 5e-5
    e = rel_to_scale(gpu.read(BD_BUF_RHS), cpu.rhs())
    say(f"rhs with ply contact maxAbs/scale = {e:.2e}")
    assert e < 1e-3


def test_two_ply_stages():
    """The solve's stages on a scene of two plies (per-layer lattices, line tables and levels) with
    their contact terms, at the reference's bars (tests/gpu/correctness.ts): Galerkin levels 1e-5,
    prolongation 1e-6, smoother sweep 1e-3, V-cycle 5e-3, PCG 2e-2, line-search energies 1e-2, and
    the FP64 true residual of a PCG to rel-res 1e-6."""
    from bd.solver import GpuSolver
    cfg = config_plies()
    S, params = cfg["scene"], cfg["params"]
    x, _ = pressed(S, step=0.6e-3)
    x[:, :2] *= np.float32(1.002)
    gpu = GpuSolver(S, params)
    gpu.set_pcg(2, 0.0, 100)
    gpu.set_state(x.astype(np.float64))
    gpu.encode_predict()
    gpu.encode_contact_detect()
    gpu.encode_project_start()
    gpu.encode_contact_terms()
    gpu.encode_assemble(True)
    cpu = O.Oracle(S, params)
    cpu.x[:] = x.astype(np.float64).ravel()
    cpu.predict()
    cpu.dx[:] = gpu.dx.ravel()
    cpu.set_ext(gpu.ext().astype(np.float64))
    cpu.set_cmeta(gpu.cmeta().astype(np.float64))
    cpu.assemble()
    cpu.load_fine_operator(gpu.stencil(0), gpu.fixed(0))
    nlev = cpu.level_count()
    say(f"two plies: {S.node_count} nodes, levels {gpu.stats()['level_nodes']}")
    assert nlev == len(gpu.stats()["level_nodes"]) and nlev >= 2
    for k in range(nlev - 1):
        _, pc = cpu.prolongation(k)
        assert np.max(np.abs(gpu.prolongation(k) - pc)) < 1e-6, f"prolongation weights level {k}"
    for k in range(1, nlev):
        assert np.array_equal(gpu.fixed(k), cpu.fixed(k)), f"fixed flags level {k}"
        Ak, Ck = gpu.stencil(k), cpu.stencil(k)
        e = rel_to_scale(Ak.reshape(len(Ak), -1), Ck.reshape(len(Ck), -1))
        say(f"galerkin level {k} maxAbs/scale = {e:.2e}")
        assert e < 1e-5
    b = gpu.read(BD_BUF_RHS)
    N = S.node_count
    gpu.write(BD_BUF_LEVEL_B, b, 0)
    gpu.write(BD_BUF_LEVEL_X, np.zeros(3 * N), 0)
    gpu.encode_smooth(0, False)
    e = rel_to_scale(gpu.read(BD_BUF_LEVEL_X, 0), cpu.smooth(0, np.zeros(3 * N), b, False))
    say(f"smoother sweep maxAbs/scale = {e:.2e}")
    assert e < 1e-3
    gpu.write(BD_BUF_LEVEL_B, b, 0)
    gpu.encode_vcycle()
    e = rel_to_scale(gpu.read(BD_BUF_LEVEL_X, 0), cpu.vcycle(b))
    say(f"V-cycle maxAbs/scale = {e:.2e}")
    assert e < 5e-3
    gpu.write(BD_BUF_RHS, b)
    gpu.encode_pcg()
    dg = gpu.read(BD_BUF_DELTA)
    e = rel_to_scale(dg, cpu.pcg(b, 2))
    say(f"PCG (2 iterations) maxAbs/scale = {e:.2e}")
    assert e < 2e-2
    gpu.encode_line_search()
    eg, mg = gpu.energy()
    ec, mc = cpu.energy_change(dg)
    err = np.max(np.abs(eg - ec)) / np.max(np.abs(ec))
    say(f"line-search energies gpu {eg} cpu {ec} rel {err:.2e}")
    assert err < 1e-2
    gpu.set_pcg(2, 1e-6, 300)
    gpu.write(BD_BUF_RHS, b)
    gpu.encode_pcg()
    st = gpu.stats()
    d = gpu.read(BD_BUF_DELTA)
    fixed = np.repeat(cpu.fixed(0) == 1, 3)
    bb = b.copy()
    bb[fixed] = 0
    r = bb - cpu.apply(0, d)
    r[fixed] = 0
    rel = np.linalg.norm(r) / np.linalg.norm(bb)
    say(f"pcg to 1e-6: {st['pcg_iterations'][0]} iterations, reported {st['rel_residual'][0]:.2e}, recomputed {rel:.2e}")
    assert st["rel_residual"][0] <= 1e-6 and rel < 1.05e-6


def oracle_ply_substep(cpu, S, params):
    """One FP64 oracle substep with the restatement's detection and terms in the order of
    encodeSubstep: predict, detect, projectStart, per Newton iteration terms then the step."""
    cpu.predict()
    x32 = cpu.x.reshape(-1, 3).astype(np.float32)
    T = O.PlyTwin(S, x32, cpu.dx.reshape(-1, 3).astype(np.float32))
    _, _, pairs, cmeta = T.detect()
    cpu.set_cmeta(cmeta.astype(np.float64))
    cpu.project_start()
    for it in range(params.newton_iterations):
        T = O.PlyTwin(S, x32, cpu.dx.reshape(-1, 3).astype(np.float32))
        cpu.set_ext(T.terms(pairs, np.float32(params.h)).astype(np.float64))
        cpu.newton_step(it == 0)
    cpu.finalize()


def ply_gap(S, x):
    """Per upper-ply node over the lower ply's interior: its height above the nearest lower node."""
    L0, L1 = S.layers[0], S.layers[1]
    lower = x[:L0.node_count]
    upper = x[L1.node_offset:L1.node_offset + L1.node_count]
    inner = (np.abs(upper[:, 0]) < 6e-3) & (np.abs(upper[:, 1]) < 7e-3)
    g = []
    for p in upper[inner]:
        k = np.argmin(np.linalg.norm(lower[:, :2] - p[:2], axis=1))
        g.append(p[2] - lower[k, 2])
    return np.array(g)


def test_upper_ply_pressed_onto_the_lower_one():
    """A centre grip presses the upper ply 1.5 mm down over 15 substeps (it starts 0.75 mm above
    the lower ply): the GPU follows the oracle over 4 substeps to 0.25 of the motion and the upper
    ply stays near the 0.7 mm contact distance."""
    from bd.solver import GpuSolver
    cfg = config_plies()
    S, params = cfg["scene"], cfg["params"]
    grips = [dict(g) for g in cfg["grips"]]
    gpu = GpuSolver(S, params, grips=grips)
    cpu = O.Oracle(S, params)
    x0 = cpu.x.copy().reshape(-1, 3)
    say("gpu: upper ply pressed onto the lower one with cross-layer contact")
    parity = None
    for k in range(15):
        advance_grips(grips, cfg["velocity"])
        gpu.set_grip_goals([g["goal"] for g in grips])
        cpu.set_grips(grips)
        gpu.encode_substep()
        oracle_ply_substep(cpu, S, params)
        st = gpu.stats()
        say(f"  substep {k}: newton {st['newton_iterations']} pcg {st['pcg_iterations'][:st['newton_iterations']]} "
            f"alpha {st['alpha'][:st['newton_iterations']]}")
        if k == 3:
            xg4, xc4 = gpu.x, cpu.x.reshape(-1, 3).copy()
            parity = (np.max(np.linalg.norm(xg4 - xc4, axis=1)), np.max(np.linalg.norm(xc4 - x0, axis=1)))
    xg = gpu.x
    assert np.isfinite(xg).all()
    g = ply_gap(S, xg)
    say(f"ply gap min {g.min() * 1e3:.3f} mm median {np.median(g) * 1e3:.3f} mm ({len(g)} nodes)")
    assert len(g) > 200 and g.min() > 0.3e-3 and np.median(g) < 0.9e-3
    err, motion = parity
    say(f"gpu vs oracle after 4 substeps: max |dx| {err * 1e3:.4f} mm of motion {motion * 1e3:.3f} mm")
    assert err < 0.25 * motion

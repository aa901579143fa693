"""GPU vs FP64 oracle, stage by stage, through the C-ABI (mirrors the reference's
tests/gpu/correctness.ts:165-275 on its worn 48x40 coupon with 3x2 bundles).

Tolerances are the reference's own GPU-vs-f64 bars (relative to each quantity's max magnitude):
fine operator 5e-5, rhs 1e-3, Galerkin levels 1e-5, one smoother sweep 1e-3, V-cycle 5e-3,
PCG step 2e-2, line-search energies 1e-2, 60-substep dynamics 0.25 of the motion.
"""
import sys

import numpy as np
import pytest

import oracle as O
from bd.capi import (make_params, BD_BUF_RHS, BD_BUF_LEVEL_B, BD_BUF_LEVEL_X, BD_BUF_DELTA)
from bd.material import DENIM
from bd.scene import compile_scene, coupon

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def corr_coupon(nu=48, nv=40, **kw):  # correctness.ts:26-45
    spec = dict(bundle=[3, 2], angle=0.3, compressU=0.03, compressV=0.02, shear=0.02, collarRows=2,
                wear=dict(center=[0, 0], axis=[1, 0], radii=[0.03, 0.03], amount=1, cellSize=0, featureScale=0.01,
                          defectThreshold=2, defectDensity=0, sparseFrom=2),
                wearStrengthLoss=[0.45, 0.9], seed=7, perturbation=0.4e-3)
    spec.update(kw)
    return coupon(nu, nv, **spec)


PARAMS = make_params()  # GPU_DEFAULTS (gpu/solver.ts:61-72), self-contact off


def rel_to_scale(a, b, skip_rows=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if skip_rows is not None and len(skip_rows):
        keep = np.ones(a.shape[0], bool)
        keep[list(skip_rows)] = False
        a, b = a[keep], b[keep]
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a)))


def kink_rows(cpu, S):
    """Rows touched by a span whose strain is within 1e-6 of the axial law's kink (correctness.ts:55-68)."""
    rows = set()
    x = cpu.x.reshape(-1, 3) + cpu.dx.reshape(-1, 3)
    L = S.layers[0]
    for s in np.nonzero((cpu.seg_flags & 3) == 1)[0]:
        a = s >> 1
        b = a + 1 if s % 2 == 0 else a + L.nx
        if abs(np.linalg.norm(x[b] - x[a]) / cpu.rest[s] - 1) < 1e-6:
            rows.update((a, b))
    return rows


def copy_state(cpu, gpu):
    """Round the oracle's state to FP32 so both sides start identically (correctness.ts:146-163)."""
    xs = cpu.x.astype(np.float32).astype(np.float64)
    vs = cpu.v.astype(np.float32).astype(np.float64)
    cpu.x[:] = xs
    cpu.v[:] = vs
    gpu.set_state(xs, vs)
    gpu.set_segments(cpu.rest, cpu.intact, cpu.cut_mask, cpu.seg_flags, cpu.fail_site, cpu.tails)
    gpu.set_field(cpu.damage)
    return xs, vs


@pytest.fixture(scope="module")
def coupon_pair():
    from bd.solver import GpuSolver
    S = compile_scene([corr_coupon()])
    cpu = O.Oracle(S, PARAMS)
    for _ in range(12):
        cpu.substep()
    say("gpu: create solver on the 48x40 worn coupon")
    gpu = GpuSolver(S, PARAMS)
    copy_state(cpu, gpu)
    return S, cpu, gpu


def test_assemble_and_hierarchy(coupon_pair):
    S, cpu, gpu = coupon_pair
    copy_state(cpu, gpu)
    cpu.predict()
    cpu.assemble()
    say("gpu: predict + assemble")
    gpu.encode_predict()
    gpu.encode_assemble(True)
    Ag = gpu.stencil(0)
    Ac = cpu.stencil(0)
    skip = kink_rows(cpu, S)
    e = rel_to_scale(Ag.reshape(len(Ag), -1), Ac.reshape(len(Ac), -1), skip)
    say(f"fine operator maxAbs/scale = {e:.2e}")
    assert e < 5e-5
    bg = gpu.read(BD_BUF_RHS)
    bc = cpu.rhs()
    e = rel_to_scale(bg, bc)
    say(f"newton rhs maxAbs/scale = {e:.2e}")
    assert e < 1e-3
    assert np.array_equal(gpu.fixed(0), cpu.fixed(0))
    # Galerkin levels from the GPU fine operator (isolates RAP, correctness.ts:199-206)
    cpu.load_fine_operator(Ag, gpu.fixed(0))
    assert cpu.level_count() == len(gpu.stats()["level_nodes"])
    for k in range(cpu.level_count() - 1):
        pg = gpu.prolongation(k)
        _, pc = cpu.prolongation(k)
        assert np.max(np.abs(pg - pc)) < 1e-6, f"prolongation weights level {k}"
    for k in range(1, cpu.level_count()):
        Ak = gpu.stencil(k)
        Ck = cpu.stencil(k)
        assert np.array_equal(gpu.fixed(k), cpu.fixed(k)), f"fixed flags level {k}"
        e = rel_to_scale(Ak.reshape(len(Ak), -1), Ck.reshape(len(Ck), -1))
        say(f"galerkin level {k} maxAbs/scale = {e:.2e}")
        assert e < 1e-5


def test_smoother_vcycle_pcg(coupon_pair):
    S, cpu, gpu = coupon_pair
    copy_state(cpu, gpu)
    cpu.predict()
    cpu.assemble()
    gpu.encode_predict()
    gpu.encode_assemble(True)
    cpu.load_fine_operator(gpu.stencil(0), gpu.fixed(0))
    b = gpu.read(BD_BUF_RHS)
    N = S.node_count
    # one forward sweep from zero (correctness.ts:464 tolerance 1e-3)
    say("gpu: one smoother sweep")
    gpu.write(BD_BUF_LEVEL_B, b, 0)
    gpu.write(BD_BUF_LEVEL_X, np.zeros(3 * N), 0)
    gpu.encode_smooth(0, False)
    xg = gpu.read(BD_BUF_LEVEL_X, 0)
    xc = cpu.smooth(0, np.zeros(3 * N), b, False)
    e = rel_to_scale(xg, xc)
    say(f"smoother sweep maxAbs/scale = {e:.2e}")
    assert e < 1e-3
    say("gpu: V-cycle")
    gpu.write(BD_BUF_LEVEL_B, b, 0)
    gpu.encode_vcycle()
    zg = gpu.read(BD_BUF_LEVEL_X, 0)
    zc = cpu.vcycle(b)
    e = rel_to_scale(zg, zc)
    say(f"V-cycle maxAbs/scale = {e:.2e}")
    assert e < 5e-3
    say("gpu: PCG (2 iterations)")
    gpu.write(BD_BUF_RHS, b)
    gpu.encode_pcg()
    dg = gpu.read(BD_BUF_DELTA)
    dc = cpu.pcg(b, PARAMS.pcg_iterations)
    e = rel_to_scale(dg, dc)
    say(f"PCG step maxAbs/scale = {e:.2e}")
    assert e < 2e-2
    # line-search energy changes along the GPU direction (correctness.ts:223-231)
    say("gpu: line search")
    gpu.encode_line_search()
    eg, mg = gpu.energy()
    ec, mc = cpu.energy_change(dg)
    err = np.max(np.abs(eg - ec)) / np.max(np.abs(ec))
    say(f"line-search energies gpu {eg} cpu {ec} rel {err:.2e}")
    assert err < 1e-2
    assert np.max(np.abs(mg - mc) / np.abs(mc)) < 1e-2


def test_pcg_relative_tolerance(coupon_pair):
    """Tolerance mode: the FP64-accumulated residual the GPU reports is the true residual of its
    FP32 iterate, recomputed here in FP64 with the oracle's operator."""
    S, cpu, gpu = coupon_pair
    copy_state(cpu, gpu)
    cpu.predict()
    cpu.assemble()
    gpu.encode_predict()
    gpu.encode_assemble(True)
    cpu.load_fine_operator(gpu.stencil(0), gpu.fixed(0))
    gpu.set_pcg(2, 1e-6, 200)
    say("gpu: PCG to rel-res 1e-6")
    gpu.encode_pcg()
    st = gpu.stats()
    d = gpu.read(BD_BUF_DELTA)
    b = gpu.read(BD_BUF_RHS)
    fixed = np.repeat(cpu.fixed(0) == 1, 3)
    b[fixed] = 0
    r = b - cpu.apply(0, d)
    r[fixed] = 0
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    say(f"pcg iterations {st['pcg_iterations'][0]} reported {st['rel_residual'][0]:.2e} recomputed {rel:.2e}")
    gpu.set_pcg(PARAMS.pcg_iterations, 0.0, 200)
    assert st["rel_residual"][0] <= 1e-6
    assert rel < 1.05e-6
    assert 1 <= st["pcg_iterations"][0] < 200


def test_pcg_convergence_check_fallback_path(coupon_pair, monkeypatch):
    """The host normally polls a pinned flag for PCG's convergence; with a poll timeout of 0 every
    check takes the synchronizing fallback (stream synchronize, then the device flag). Both paths must
    stop at the same iteration with the same bits (the kernels and their reductions are identical)."""
    from bd.solver import GpuSolver
    S, cpu, gpu = coupon_pair
    copy_state(cpu, gpu)
    monkeypatch.setenv("BD_POLL_TIMEOUT_MS", "0")
    slow = GpuSolver(S, PARAMS)
    monkeypatch.delenv("BD_POLL_TIMEOUT_MS")
    copy_state(cpu, slow)
    res = []
    for g in (gpu, slow):
        g.encode_predict()
        g.encode_assemble(True)
        g.set_pcg(2, 1e-6, 200)
        before = g.stats()["poll_fallbacks"]
        g.encode_pcg()
        st = g.stats()
        g.set_pcg(PARAMS.pcg_iterations, 0.0, 200)
        res.append((st["pcg_iterations"][0], st["rel_residual"][0], g.read(BD_BUF_DELTA), st["poll_fallbacks"] - before))
    say(f"polled: {res[0][0]} iterations, rel {res[0][1]:.3e}; fallback: {res[1][0]} iterations, rel {res[1][1]:.3e}, "
        f"{res[1][3]} synchronizing checks")
    assert res[0][3] == 0 and res[1][3] == res[1][0]
    assert res[0][0] == res[1][0] and res[0][1] == res[1][1]
    assert np.array_equal(res[0][2], res[1][2])
    del slow


def test_dynamics_60_substeps(coupon_pair):
    S, cpu, gpu = coupon_pair
    xs, _ = copy_state(cpu, gpu)
    steps = 60
    for _ in range(steps):
        cpu.substep()
    say("gpu: 60 substeps")
    for _ in range(steps):
        gpu.encode_substep()
    xg = gpu.read(0)
    moved = np.max(np.abs(cpu.x - xs))
    err = np.max(np.abs(xg - cpu.x))
    say(f"dynamics: max |x_gpu - x_cpu| = {err * 1e6:.2f} um, max motion {moved * 1e6:.1f} um")
    assert err < 0.25 * moved


def test_max_speed_counter(coupon_pair):
    """stats()["max_speed"] is the largest committed node speed since the previous read (the value a
    rest detector polls instead of reading back every velocity), and reading it clears it."""
    S, cpu, gpu = coupon_pair

    def speeds():  # the velocity buffer is read back node-major, xyz per node
        return np.sqrt((gpu.read(1).astype(np.float64).reshape(-1, 3) ** 2).sum(axis=1))

    gpu.stats()
    for _ in range(3):
        gpu.encode_substep()
    last3 = speeds().max()
    first = gpu.stats()["max_speed"]
    again = gpu.stats()["max_speed"]
    gpu.encode_substep()
    last1 = speeds().max()
    one = gpu.stats()["max_speed"]
    say(f"max speed: counter over 3 substeps {first:.6e} m/s (last substep's velocities {last3:.6e}), "
        f"after a read {again}, over 1 substep {one:.6e} (velocities {last1:.6e})")
    assert last3 > 0.0
    # over three substeps the counter is at least the last one's largest speed
    assert first >= last3 * (1 - 1e-6)
    assert again == 0.0
    assert abs(one - last1) <= 1e-5 * last1


@pytest.mark.parametrize("max_coarsest,fp16", [(400, False), (1100, False), (1100, True)])
def test_exact_coarse_solve(coupon_pair, max_coarsest, fp16):
    """BD_COARSE_EXACT: the GPU's FP64 banded Cholesky + dense inverse of the coarsest level in the
    V-cycle and PCG, against the oracle's banded Cholesky solve (bars as above). Coarsest 13 x 11
    factors in the LDS window, 25 x 21 (bandwidth 158) in place in global memory; and the 25 x 21
    exact solve under FP16 coarse-level operators (BD_PRECISION_FP16), the oracle rounding them the
    same way."""
    from bd.capi import BD_PRECISION_FP16, BD_PRECISION_FP32
    from bd.solver import GpuSolver
    S, cpu, _ = coupon_pair
    params = make_params(coarse_solve=1, max_coarsest_nodes=max_coarsest,
                         precision=BD_PRECISION_FP16 if fp16 else BD_PRECISION_FP32)
    gpu = GpuSolver(S, params)
    copy_state(cpu, gpu)
    ref = O.Oracle(S, params)
    ref.x[:] = cpu.x
    ref.v[:] = cpu.v
    ref.set_grips(None)
    say("gpu: exact coarse solve: predict + assemble")
    ref.predict()
    gpu.encode_predict()
    gpu.encode_assemble(True)
    ref.assemble()
    ref.load_fine_operator(gpu.stencil(0), gpu.fixed(0))
    b = gpu.read(BD_BUF_RHS)
    say("gpu: exact coarse solve: V-cycle")
    gpu.write(BD_BUF_LEVEL_B, b, 0)
    gpu.encode_vcycle()
    e = rel_to_scale(gpu.read(BD_BUF_LEVEL_X, 0), ref.vcycle(b))
    say(f"V-cycle (exact coarsest) maxAbs/scale = {e:.2e}")
    assert e < 5e-3
    say("gpu: exact coarse solve: PCG (2 iterations)")
    gpu.write(BD_BUF_RHS, b)
    gpu.encode_pcg()
    e = rel_to_scale(gpu.read(BD_BUF_DELTA), ref.pcg(b, params.pcg_iterations))
    say(f"PCG step (exact coarsest) maxAbs/scale = {e:.2e}")
    assert e < 2e-2


def test_lagged_exact_coarse_solve(coupon_pair):
    """BD_COARSE_EXACT_LAGGED: the coarsest level factored once per substep, at its first Newton
    iteration, while the Galerkin levels are rebuilt at every one (lag_coarse off). The profile counts
    one factorization per substep against one per Newton iteration for BD_COARSE_EXACT; four substeps
    of 2 Newton iterations, each solved to rel-res 1e-6, follow the oracle's substeps of the same mode,
    and the lagged and the exact runs end within the same bar of each other."""
    from bd.capi import BD_COARSE_EXACT, BD_COARSE_EXACT_LAGGED
    from bd.solver import GpuSolver
    S = coupon_pair[0]
    out = {}
    for mode in (BD_COARSE_EXACT, BD_COARSE_EXACT_LAGGED):
        params = make_params(coarse_solve=mode, lag_coarse=False, history=False, pcg_rel_tol=1e-6,
                             pcg_max_iterations=300)
        ref = O.Oracle(S, params)
        gpu = GpuSolver(S, params)
        xs, _ = copy_state(ref, gpu)
        gpu.profile_enable(True)
        gpu.profile_read()
        say(f"gpu: 4 substeps, coarse_solve {mode}")
        for _ in range(4):
            ref.substep()
            gpu.encode_substep()
            st = gpu.stats()
            n = st["newton_iterations"]
            assert n == 2
            assert max(st["rel_residual"][:n]) <= 1.05e-6
        factors = gpu.profile_read().get("coarse_factor", (0, 0.0, 0.0))[0]
        xg = gpu.read(0)
        moved = np.max(np.abs(ref.x - xs))
        err = np.max(np.abs(xg - ref.x))
        say(f"  {factors} coarsest factorizations; max |x_gpu - x_oracle| = {err * 1e6:.3f} um, motion {moved * 1e6:.1f} um")
        assert err < 0.02 * moved
        out[mode] = (xg, factors, moved)
        gpu.close()
    assert out[BD_COARSE_EXACT][1] == 8 and out[BD_COARSE_EXACT_LAGGED][1] == 4
    d = np.max(np.abs(out[BD_COARSE_EXACT][0] - out[BD_COARSE_EXACT_LAGGED][0]))
    say(f"lagged against exact: {d * 1e6:.3f} um")
    assert d < 0.02 * out[BD_COARSE_EXACT][2]

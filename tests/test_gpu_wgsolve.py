"""The one-workgroup PCG (BD_PCG_WORKGROUP, csrc/wgsolve.hip) on the GPU through the C-ABI: the whole
solve of one Newton system in one workgroup, and a batch's solves in one launch.

* against the FP64 oracle on the reference's worn 48 x 40 coupon (tests/gpu/correctness.ts:165-275):
  a fixed-count PCG step within the reference's 2e-2, and a solve to rel-res 1e-6 whose true residual,
  recomputed in FP64 with the same operator, is at most 1.05e-6;
* against the grid path: the same preconditioned CG with its FP64 sums in another order, so on the
  64 x 64 sheet (cfg 1), with the exact coarsest solve, and on the fold with self-contact the PCG
  iteration counts agree within a few percent and the trajectories within 0.03 of the motion (the
  reference's converged-schedule bar, correctness.ts:263-271);
* a fused batch of uneven patches (each Newton iteration's solves of all patches in one launch, one
  workgroup per patch) equals each patch's own one-workgroup substeps bit for bit;
* the limits are refused with their reason (several layers, FP16 operators).
"""
import sys

import numpy as np
import pytest

import oracle as O
from bd.capi import BD_BUF_DELTA, BD_BUF_RHS, BD_PCG_WORKGROUP, BD_PRECISION_FP16, make_params
from bd.configs import advance_grips, config1, config_fold, config_patches, config_plies, config_sheet

from test_gpu_parity import PARAMS, copy_state, corr_coupon, rel_to_scale

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def coupon_wg():
    from bd.scene import compile_scene
    from bd.solver import GpuSolver
    S = compile_scene([corr_coupon()])
    cpu = O.Oracle(S, PARAMS)
    for _ in range(12):
        cpu.substep()
    say("gpu: worn coupon, one-workgroup PCG")
    gpu = GpuSolver(S, PARAMS)
    copy_state(cpu, gpu)
    gpu.set_pcg_mode(BD_PCG_WORKGROUP)
    return S, cpu, gpu


def assembled(cpu, gpu):
    copy_state(cpu, gpu)
    cpu.predict()
    cpu.assemble()
    gpu.encode_predict()
    gpu.encode_assemble(True)
    cpu.load_fine_operator(gpu.stencil(0), gpu.fixed(0))
    return gpu.read(BD_BUF_RHS)


def test_fixed_count_step_against_oracle(coupon_wg):
    S, cpu, gpu = coupon_wg
    b = assembled(cpu, gpu)
    gpu.encode_pcg()
    dg = gpu.read(BD_BUF_DELTA)
    dc = cpu.pcg(b, PARAMS.pcg_iterations)
    e = rel_to_scale(dg, dc)
    say(f"one-workgroup PCG ({PARAMS.pcg_iterations} iterations) maxAbs/scale = {e:.2e}")
    assert e < 2e-2


def test_tolerance_solve_true_residual(coupon_wg):
    S, cpu, gpu = coupon_wg
    assembled(cpu, gpu)
    gpu.set_pcg(2, 1e-6, 200)
    gpu.encode_pcg()
    st = gpu.stats()
    d = gpu.read(BD_BUF_DELTA)
    b = gpu.read(BD_BUF_RHS)
    fixed = np.repeat(cpu.fixed(0) == 1, 3)
    b[fixed] = 0
    r = b - cpu.apply(0, d)
    r[fixed] = 0
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    gpu.set_pcg(PARAMS.pcg_iterations, 0.0, 200)
    say(f"one-workgroup PCG to 1e-6: {st['pcg_iterations'][0]} iterations, reported {st['rel_residual'][0]:.2e}, "
        f"FP64 recomputed {rel:.2e}")
    assert st["rel_residual"][0] <= 1e-6
    assert rel < 1.05e-6


def trajectories(cfg, frames, floor=None):
    from bd.solver import GpuSolver
    S, P = cfg["scene"], cfg["params"]
    out = []
    for mode in (0, BD_PCG_WORKGROUP):
        grips = [dict(g) for g in cfg.get("grips", [])]
        g = GpuSolver(S, P, grips or None)
        if floor is not None:
            g.set_floor(floor)
        g.set_pcg_mode(mode)
        x0 = g.read(0)
        its = 0
        for _ in range(frames):
            if grips:
                advance_grips(grips, cfg["velocity"])
                g.set_grip_goals(np.array([q["goal"] for q in grips]))
            g.encode_substep()
            st = g.stats()
            its += sum(st["pcg_iterations"][:st["newton_iterations"]])
            assert np.isfinite(g.read(0)).all()
        out.append((g.read(0), its, x0))
        g.close()
    (xg, ig, x0), (xw, iw, _) = out
    moved = np.max(np.abs(xg - x0))
    err = np.max(np.abs(xw - xg))
    say(f"  {frames} frames: PCG iterations grid {ig}, one workgroup {iw}; max |x_wg - x_grid| {err * 1e6:.2f} um "
        f"of {moved * 1e3:.3f} mm motion")
    assert abs(iw - ig) <= 0.05 * ig + 2
    assert err <= 0.03 * moved


def test_sheet_trajectory_against_grid():
    say("gpu: cfg 1 sheet, grid and one-workgroup PCG")
    trajectories(config1(), 20)


def test_exact_coarsest_against_grid():
    cfg = config_sheet(n=100, region=24, history=False)
    P = cfg["params"]
    P.lag_coarse = 0
    P.coarse_solve = 1
    say("gpu: 100 x 100 sheet, exact coarsest level, grid and one-workgroup PCG")
    trajectories(cfg, 6)


def test_fold_with_contact_against_grid():
    # the converged-schedule bar needs converged solves: the fold's landing takes up to 330 PCG
    # iterations per Newton solve under the plain schedule, and solves stopped at the configuration's
    # cap of 100 (rel-res up to 2e-2) leave trajectories that depend on the last rounding of each path
    cfg = config_fold()
    cfg["params"].pcg_max_iterations = 400
    say("gpu: fold with self-contact and the floor, grid and one-workgroup PCG (every solve converged)")
    trajectories(cfg, 8, floor=cfg["floor_z"])


def test_fused_batch_matches_single_solves():
    from bd.solver import Batch, GpuSolver
    cfg = config_patches(count=6, side_range=(32, 96), seed=11)
    say(f"gpu: fused batch of {len(cfg['patches'])} patches, sides {[p['side'] for p in cfg['patches']]}")
    single, batched, grips = [], [], []
    for p in cfg["patches"]:
        ga, gb = [dict(g) for g in p["grips"]], [dict(g) for g in p["grips"]]
        a, b = GpuSolver(p["scene"], p["params"], ga), GpuSolver(p["scene"], p["params"], gb)
        a.set_pcg_mode(BD_PCG_WORKGROUP)
        b.set_pcg_mode(BD_PCG_WORKGROUP)
        single.append(a)
        batched.append(b)
        grips.append(ga)
    batch = Batch(batched)
    for k in range(5):
        for p, a, b, ga in zip(cfg["patches"], single, batched, grips):
            advance_grips(ga, p["velocity"])
            goals = np.array([g["goal"] for g in ga])
            a.set_grip_goals(goals)
            b.set_grip_goals(goals)
            a.encode_substep()
        batch.substep()
        batch.sync()
        its = []
        for i, (a, b) in enumerate(zip(single, batched)):
            xa, xb = a.read(0), b.read(0)
            assert np.array_equal(xa, xb), f"frame {k} patch {i}: {np.max(np.abs(xa - xb)):.3e}"
            assert np.array_equal(a.read(1), b.read(1)), (k, i)
            sa, sb = a.stats(), b.stats()
            n = sa["newton_iterations"]
            assert list(sa["pcg_iterations"][:n]) == list(sb["pcg_iterations"][:n])
            assert max(sb["rel_residual"][:n]) <= 1.05e-6, (k, i, sb)
            its.append(sum(sb["pcg_iterations"][:n]))
        say(f"  frame {k}: PCG iterations per patch {its}, bit-exact")
    batched[0].set_timestep(1 / 240)
    with pytest.raises(RuntimeError, match="changed"):
        batch.substep()
    batch.close()
    for g in single + batched:
        g.close()


def test_refusals():
    from bd.solver import GpuSolver
    cfg = config_plies()
    g = GpuSolver(cfg["scene"], cfg["params"])
    with pytest.raises(RuntimeError, match="several layers"):
        g.set_pcg_mode(BD_PCG_WORKGROUP)
    g.close()
    c = config_sheet(n=136, region=16, history=False)
    P = c["params"]
    P.precision = BD_PRECISION_FP16
    h = GpuSolver(c["scene"], P, c["grips"])
    with pytest.raises(RuntimeError, match="FP16"):
        h.set_pcg_mode(BD_PCG_WORKGROUP)
    h.close()
    with pytest.raises(RuntimeError, match="unknown mode"):
        GpuSolver(c["scene"], make_params()).set_pcg_mode(7)

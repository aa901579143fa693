"""The workload configurations end to end on the GPU, through the C-ABI, against the FP64 oracle and
the FP32 contact restatement.

* cfg 1, the plain method on a 64 x 64 sheet (0.5 m, one corner grip moved +x 2 mm/step, 120 steps,
  3 Newton iterations, 3-level MG-PCG to rel-res 1e-5): GPU and oracle run the same 120 frames from the
  same FP32 state and their positions and velocities are compared at every frame;
* the reference's converged-schedule dynamics check (tests/gpu/correctness.ts:263-271): 5 Newton x 3 PCG
  iterations on the 48 x 40 coupon, 60 substeps, within 0.03 of the motion;
* cfg 2, the fold at scale (256 x 256 with a 64 x 64 breakable region across the fold, self-contact,
  two grips pulled apart): for the first 24 frames, every frame's contact detection (cells, records,
  candidate faces, bounds) is compared bit for bit with the FP32 restatement on the GPU's own state,
  and its damage field and slip (codes, cut masks, flags, intact fractions, tails, fail sites, rest
  lengths, rupture and pull-out counts) bit for bit with the FP64 oracle on the GPU's own committed
  positions; all 240 frames stay finite with every Newton solve at rel-res <= 1.05e-6;
* cfg 3, the same scene with the mixed-precision method: 60 frames at 256 x 256 with every Newton
  solve at rel-res <= 1.05e-6; and, on the same scene at 128 x 128 (a 32 x 32 region; the FP64
  oracle needs about a minute per substep at 256 x 256), one substep from the GPU's state at sampled
  frames against the FP64 oracle's substep driven by the same contact restatement (detection on the
  GPU's own predicted step): the position error is reported per frame and held under 1e-4 m once the
  fold has landed; in the first frame, where the upper layer is in free fall onto the lower one and 3
  Newton iterations leave the step unconverged, a different line-search step at one iteration moves
  single nodes by up to ~0.15 mm, so that frame is held to 5 % of its motion.
"""
import sys

import numpy as np
import pytest

import oracle as O
from bd.capi import BD_BUF_DX, make_params
from bd.configs import advance_grips, config1, config2, led_goals
from bd.scene import compile_scene, coupon

from test_gpu_contact import f32_gnorm, sorted_records
from test_gpu_history import compare_state

pytestmark = pytest.mark.gpu
K = 4


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_cfg1_positions_and_velocities_every_frame():
    """Each of the 120 frames twice: the oracle restarted from the GPU's state at the frame's start
    (one substep apart, positions within 1e-3 of the frame's motion and velocities within 1e-3 of
    the largest speed), and a second oracle running the whole trajectory on its own (within the
    reference's production bar, 0.25 of the motion, correctness.ts:233)."""
    from bd.solver import GpuSolver
    cfg = config1()
    S, P = cfg["scene"], cfg["params"]
    grips = [dict(g) for g in cfg["grips"]]
    gpu = GpuSolver(S, P, grips)
    cpu = O.Oracle(S, P, grips)
    free = O.Oracle(S, P, grips)
    x0 = S.position.astype(np.float32).astype(np.float64).ravel()
    gpu.set_state(x0)
    free.x[:] = x0
    say("gpu: cfg 1, 64 x 64 sheet, 120 frames")
    worst_x = worst_v = worst_t = 0.0
    for k in range(cfg["steps"]):
        advance_grips(grips, cfg["velocity"])
        cpu.set_grips(grips)
        free.set_grips(grips)
        gpu.set_grip_goals(np.array([g["goal"] for g in grips]))
        xs, vs = gpu.read(0), gpu.read(1)
        cpu.x[:] = xs
        cpu.v[:] = vs
        gpu.encode_substep()
        cpu.substep()
        free.substep()
        xg, vg = gpu.read(0), gpu.read(1)
        assert np.isfinite(xg).all() and np.isfinite(vg).all()
        step = np.max(np.abs(cpu.x - xs))
        ex = np.max(np.abs(xg - cpu.x)) / step
        ev = np.max(np.abs(vg - cpu.v)) / np.max(np.abs(cpu.v))
        et = np.max(np.abs(xg - free.x)) / np.max(np.abs(free.x - x0))
        worst_x, worst_v, worst_t = max(worst_x, ex), max(worst_v, ev), max(worst_t, et)
        if k % 20 == 19 or k == 0:
            st = gpu.stats()
            say(f"  frame {k + 1}: one step: |x_gpu - x_fp64| / frame motion {ex:.2e}, |v_gpu - v_fp64| / max|v| "
                f"{ev:.2e}; trajectory {et:.2e} of the motion (pcg {st['pcg_iterations'][:st['newton_iterations']]})")
        assert ex < 1e-3, (k, ex)
        assert ev < 1e-3, (k, ev)
        assert et < 0.25, (k, et)
    say(f"cfg 1 worst over 120 frames: one step {worst_x:.2e} of the frame motion, velocities {worst_v:.2e} of "
        f"max |v|, trajectory {worst_t:.2e} of the motion")


def test_converged_schedule_dynamics():  # correctness.ts:263-271
    from bd.solver import GpuSolver
    from bd.material import DENIM
    S = compile_scene([coupon(48, 40, material=DENIM)])
    P = make_params(newton_iterations=5, newton_max=5, pcg_iterations=3)
    cpu = O.Oracle(S, P)
    for _ in range(12):
        cpu.substep()
    gpu = GpuSolver(S, P)
    xs = cpu.x.astype(np.float32).astype(np.float64)
    vs = cpu.v.astype(np.float32).astype(np.float64)
    cpu.x[:] = xs
    cpu.v[:] = vs
    gpu.set_state(xs, vs)
    gpu.set_segments(cpu.rest, cpu.intact, cpu.cut_mask, cpu.seg_flags, cpu.fail_site, cpu.tails)
    gpu.set_field(cpu.damage)
    for _ in range(60):
        cpu.substep()
    say("gpu: converged schedule, 60 substeps")
    for _ in range(60):
        gpu.encode_substep()
    xg = gpu.read(0)
    moved = np.max(np.abs(cpu.x - xs))
    err = np.max(np.abs(xg - cpu.x))
    say(f"converged schedule: max |x_gpu - x_cpu| = {err * 1e6:.2f} um, max motion {moved * 1e6:.1f} um")
    assert err < 0.03 * moved


def oracle_substep_on_gpu_detection(cpu, S, params, dx_pred):
    """One FP64 oracle substep whose contact detection sees the GPU's own FP32 predicted step (so both
    solve with the same candidate faces and bounds: detection is bit-exact on equal inputs), with the
    FP32 restatement's terms between its Newton iterations at the oracle's own iterates."""
    g = f32_gnorm(params.gravity)
    cpu.predict()
    x32 = cpu.x.reshape(-1, 3).astype(np.float32)
    T = O.ContactTwin(S, x32, dx_pred.astype(np.float32), flags=cpu.seg_flags)
    _, _, pairs, cmeta = T.detect()
    cpu.set_cmeta(cmeta.astype(np.float64))
    cpu.set_floor(None)
    cpu.project_start()
    for it in range(params.newton_iterations):
        T = O.ContactTwin(S, x32, cpu.dx.reshape(-1, 3).astype(np.float32), flags=cpu.seg_flags)
        _, ext = T.terms(pairs, cmeta, np.float32(params.h), g)
        cpu.set_ext(ext.astype(np.float64))
        cpu.set_cmeta(cmeta.astype(np.float64))
        cpu.newton_step(it == 0)
    cpu.finalize()


def grip_centroids(x, grips):
    return [np.average(x[g["nodes"]], axis=0, weights=g["weights"]) for g in grips]


def sync_oracle(cpu, gpu):
    """The oracle takes the GPU's state: positions, velocities, spans, tails and the damage field."""
    cpu.x[:] = gpu.read(0)
    cpu.v[:] = gpu.read(1)
    cpu.rest[:] = gpu.read(12)
    cpu.intact[:] = gpu.read(13)
    cpu.cut_mask[:] = gpu.read(14)
    cpu.seg_flags[:] = gpu.read(15)
    cpu.fail_site[:] = gpu.read(16)
    cpu.tails[:] = gpu.read(17)
    cpu.damage[:] = gpu.read(18)


def test_cfg2_fold_contact_and_damage_bit_exact_per_frame():
    from bd.solver import GpuSolver
    cfg = config2()
    S, P = cfg["scene"], cfg["params"]
    grips = [dict(g) for g in cfg["grips"]]
    gpu = GpuSolver(S, P, grips)
    cpu = O.Oracle(S, P)
    goals = np.array([g["goal"] for g in grips])
    checked = 24
    say(f"gpu: cfg 2, {S.node_count}-node fold with self-contact, {cfg['steps']} frames, the first {checked} checked bit for bit")
    candidates = bounded = ruptures = worst = 0
    for k in range(cfg["steps"]):
        x = gpu.x
        goals = led_goals(goals, cfg["velocity"], grip_centroids(x, grips), cfg["max_lead"])
        gpu.set_grip_goals(goals)
        if k < checked:
            # detection on the GPU's own state: start positions and the predicted step
            gpu.encode_predict()
            dx = gpu.vec(BD_BUF_DX).astype(np.float32)
            gpu.encode_contact_detect()
            T = O.ContactTwin(S, x.astype(np.float32), dx, flags=gpu.read(15))
            cells, items, pairs, cmeta = T.detect()
            gcells = gpu.cells()
            assert np.array_equal(gcells, cells), f"frame {k}: cell starts / counts"
            rg, rc = sorted_records(gcells, gpu.items()), sorted_records(cells, items)
            assert rg.keys() == rc.keys(), f"frame {k}: occupied cells"
            for h in rc:
                assert np.array_equal(rg[h], rc[h]), f"frame {k}: records of cell {h}"
            assert np.array_equal(gpu.pairs(), pairs), f"frame {k}: candidates, keys, row marks"
            assert np.array_equal(gpu.cmeta().view(np.uint32), cmeta.view(np.uint32)), f"frame {k}: bounds"
            candidates += int((pairs[:, :K] != 0xFFFFFFFF).sum())
            bounded += int((cmeta[:, 0, 3] < 1e8).sum())
            # the substep without history, then the damage field and slip on both sides from the
            # GPU's committed state
            gpu.set_history_enabled(False)
            gpu.encode_substep()
            gpu.set_history_enabled(True)
            sync_oracle(cpu, gpu)
            r0 = gpu.stats()["ruptures"]
            b0 = cpu.broken_count
            gpu.encode_damage()
            cpu.update_damage()
            compare_state(gpu, cpu, f"frame {k} damage")
            gpu.encode_slip()
            cpu.update_slip()
            compare_state(gpu, cpu, f"frame {k} slip")
            assert gpu.stats()["ruptures"] - r0 == cpu.broken_count - b0
            ruptures += cpu.broken_count - b0
        else:
            gpu.encode_substep()
        st = gpu.stats()
        rel = max(st["rel_residual"][:st["newton_iterations"]])
        worst = max(worst, rel)
        if k % 24 == 23:
            say(f"  frame {k + 1}: pcg {st['pcg_iterations'][:st['newton_iterations']]}, rel-res {rel:.2e}, "
                f"ruptures {st['ruptures']}")
        assert rel <= 1.05e-6, (k, st)
        assert np.isfinite(gpu.x).all(), k
    damaged = int(np.count_nonzero((gpu.read(18) > 0)))
    say(f"cfg 2: {candidates} candidate faces and {bounded} bounded nodes checked over {checked} frames, "
        f"{ruptures} ruptures in them; {gpu.stats()['ruptures']} ruptures and {damaged} damaged crossings after "
        f"{cfg['steps']}; worst rel-res {worst:.2e}")
    assert candidates > 0 and bounded > 0


def test_cfg3_mixed_precision_at_scale():
    from bd.solver import GpuSolver
    cfg = config2(mixed=True)
    S, P = cfg["scene"], cfg["params"]
    grips = [dict(g) for g in cfg["grips"]]
    gpu = GpuSolver(S, P, grips)
    goals = np.array([g["goal"] for g in grips])
    say(f"gpu: cfg 3 (mixed precision), {S.node_count}-node fold, 60 frames")
    iters = 0
    for k in range(60):
        goals = led_goals(goals, cfg["velocity"], grip_centroids(gpu.x, grips), cfg["max_lead"])
        gpu.set_grip_goals(goals)
        gpu.encode_substep()
        st = gpu.stats()
        iters += sum(st["pcg_iterations"][:st["newton_iterations"]])
        assert max(st["rel_residual"][:st["newton_iterations"]]) <= 1.05e-6, (k, st)
        assert np.isfinite(gpu.x).all(), k
    say(f"cfg 3: {iters / 60:.1f} PCG iterations per frame over 60 frames")


def test_cfg3_mixed_positions_against_fp64_oracle():
    from bd.solver import GpuSolver
    cfg = config2(n=128, region=32, mixed=True)
    S, P = cfg["scene"], cfg["params"]
    grips = [dict(g) for g in cfg["grips"]]
    gpu = GpuSolver(S, P, grips)
    # the oracle solves each Newton system to the same tolerance in FP64 (exact coarsest level, FP64
    # operators); its contact terms come from the FP32 restatement as the GPU's do
    cpu = O.Oracle(S, make_params(h=P.h, newton_iterations=3, pcg_rel_tol=1e-6, pcg_max_iterations=400,
                                  max_coarsest_nodes=1100, coarse_sweeps=2, lag_coarse=False, line_search=True,
                                  history=False, self_contact=True, coarse_solve=1))
    goals = np.array([g["goal"] for g in grips])
    sample = {0, 12, 24}
    errs = []
    say(f"gpu: cfg 3 scene at 128 x 128 (mixed precision), samples at frames {sorted(sample)}")
    for k in range(max(sample) + 1):
        x = gpu.x
        goals = led_goals(goals, cfg["velocity"], grip_centroids(x, grips), cfg["max_lead"])
        gpu.set_grip_goals(goals)
        if k in sample:
            sync_oracle(cpu, gpu)
            for g, goal in zip(grips, goals):
                g["goal"] = goal
            cpu.set_grips(grips)
            gpu.encode_predict()
            oracle_substep_on_gpu_detection(cpu, S, cpu.params, gpu.vec(BD_BUF_DX))
        gpu.encode_substep()
        st = gpu.stats()
        assert max(st["rel_residual"][:st["newton_iterations"]]) <= 1.05e-6, (k, st)
        if k in sample:
            xg = gpu.read(0)
            e = float(np.max(np.abs(xg - cpu.x)))
            motion = float(np.max(np.abs(cpu.x - x.ravel())))
            errs.append(e)
            say(f"  frame {k}: max |x_mixed - x_fp64| = {e * 1e6:.3f} um (frame motion {motion * 1e3:.3f} mm), "
                f"pcg gpu {st['pcg_iterations'][:st['newton_iterations']]} alpha {st['alpha'][:st['newton_iterations']]}, "
                f"oracle {cpu.pcg_iterations} (last alpha {cpu.last_alpha})")
            assert e < (0.05 * motion if k == 0 else 1e-4)
    say(f"cfg 3: position error against the FP64 oracle per sampled frame {['%.3f um' % (v * 1e6) for v in errs]}")

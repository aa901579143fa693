"""The learned coarse-space correction on the GPU (csrc/learned.hip) against the FP64 restatement
(oracle/learned_oracle.py), and what it gains.

* forward: the fields Z the GPU predicts from a level operator equal the oracle's (bf16 rounding
  emulated at the GPU's storage points) within 2e-3 of their largest entry;
* loss and weight gradient of a recorded sample equal the oracle's (same rounding) within 1e-3
  (loss) and 2e-2 (gradient, relative to its norm; cosine above 0.999); the oracle's gradient itself
  is pinned by finite differences in tests/test_learned_oracle.py;
* training on samples that plain solves record lowers the loss on held-out samples (recorded by the
  plain method on later frames: plain MG-PCG judges the net's error); with the trained fields the
  deflated solve needs no more PCG iterations than the same method without the net (within 5 %,
  no fallback). On the hanging 256 x 256 sheet the measured gain is small: the fields capture most
  of the Newton directions' energy but little of the slow modes' (the printed losses and counts);
* a net whose fields are not finite falls back to plain MG-PCG: same iteration count, a fallback
  counted, the solve still reaching its tolerance.
"""
import sys

import numpy as np
import pytest

import learned_oracle as LO
from bd.configs import advance_grips, config_sheet

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def frames(gpu, cfg, goals, n):
    it = []
    for _ in range(n):
        goals = goals + np.array(cfg["velocity"])
        gpu.set_grip_goals(goals)
        gpu.encode_substep()
        st = gpu.stats()
        assert max(st["rel_residual"][:st["newton_iterations"]]) <= 1.05e-6, st
        it.append(sum(st["pcg_iterations"][:st["newton_iterations"]]))
    return goals, it


def level_shape(gpu):
    info = gpu.learn_info()
    nodes = info["nodes"]
    n = int(round(np.sqrt(nodes)))
    assert n * n == nodes
    return info, n


@pytest.fixture(scope="module")
def recorded():
    from bd.solver import GpuSolver
    cfg = config_sheet(n=65, region=16, history=False)
    P = cfg["params"]
    P.lag_coarse = 0
    gpu = GpuSolver(cfg["scene"], P, cfg["grips"])
    gpu.learn_record(6)
    goals = np.array([g["goal"] for g in cfg["grips"]])
    say("gpu: 65 x 65 sheet, 2 plain frames recording samples")
    frames(gpu, cfg, goals, 2)
    info, n = level_shape(gpu)
    say(f"net level {info['level']} ({n} x {n}), {info['samples']} samples")
    assert info["samples"] == 6
    return gpu, n


def test_forward_matches_oracle(recorded):
    gpu, n = recorded
    b, e, op, fx = gpu.learn_sample(5)  # recorded at the last Newton iteration: the current hierarchy
    gpu.set_learned(True)
    z = gpu.learn_fields()
    gpu.set_learned(False)
    w = gpu.learn_weights().astype(np.float64)
    X = LO.features(op, fx, n, n, b[0].T[:n * n])
    zo, _ = LO.forward(w, X, n, n, fx, emulate_bf16=True)
    zg = np.concatenate([z[k].T[:n * n] for k in range(LO.K_FIELDS)], 1)
    err = np.max(np.abs(zg - zo)) / np.max(np.abs(zo))
    say(f"forward: max |z_gpu - z_oracle| / max |z| = {err:.2e}")
    assert err < 2e-3


def test_loss_and_gradient_match_oracle(recorded):
    gpu, n = recorded
    for sample in (0, 4):
        b, e, op, fx = gpu.learn_sample(sample)
        w = gpu.learn_weights().astype(np.float64)
        lg, gg = gpu.learn_loss_grad(sample)
        targets = [b[t].T[:n * n].astype(np.float64) for t in range(2)]
        lo, go = LO.loss_grad(w, op, fx, n, n, targets, e, emulate_bf16=True)
        rel = np.linalg.norm(gg - go) / np.linalg.norm(go)
        cos = float(gg @ go / (np.linalg.norm(gg) * np.linalg.norm(go)))
        say(f"sample {sample}: energies {e}, loss gpu {lg:.6f} oracle {lo:.6f}; gradient rel {rel:.2e} cos {cos:.6f}")
        assert abs(lg - lo) <= 1e-3 * max(abs(lo), 1e-3)
        assert rel < 2e-2 and cos > 0.999
        for li, (i, o, wo, bo) in enumerate(LO.LAYERS):
            seg = slice(wo, bo + o)
            assert np.linalg.norm(gg[seg] - go[seg]) <= 5e-2 * np.linalg.norm(go[seg]) + 1e-12, li


def test_training_lowers_held_out_loss_without_costing_iterations():
    from bd.solver import GpuSolver
    cfg = config_sheet(n=256, region=64, history=False)
    P = cfg["params"]
    P.lag_coarse = 0
    P.coarse_solve = 1
    P.max_coarsest_nodes = 1100
    goals = np.array([g["goal"] for g in cfg["grips"]])
    train_frames, eval_frames = 12, 8
    gpu = GpuSolver(cfg["scene"], P, cfg["grips"])
    gpu.learn_record(3 * train_frames)
    say(f"gpu: 256 x 256 sheet, {train_frames} plain frames recording samples")
    goals, it_train = frames(gpu, cfg, goals, train_frames)
    x, v = gpu.read(0), gpu.read(1)
    losses = gpu.learn_train(200, lr=3e-3, batch=3)
    say(f"training loss by epoch: {losses[0]:.4f} -> {losses[len(losses) // 2]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.8 * losses[0]
    w = gpu.learn_weights()
    gpu.close()

    def run(weights, learned):
        g = GpuSolver(cfg["scene"], P, cfg["grips"])
        g.set_state(x, v)
        if weights is not None:
            g.set_learn_weights(weights)
        g.set_learned(learned)
        if not learned:
            g.learn_record(3 * eval_frames)
        _, it = frames(g, cfg, goals.copy(), eval_frames)
        st = g.stats()
        return g, sum(it), st

    plain, it_plain, _ = run(None, False)
    held_out = [plain.learn_loss_grad(q)[0] for q in range(plain.learn_info()["samples"])]
    plain.set_learn_weights(w)
    held_trained = [plain.learn_loss_grad(q)[0] for q in range(plain.learn_info()["samples"])]
    plain.close()
    untrained, it_untrained, _ = run(None, True)
    untrained.close()
    trained, it_trained, st = run(w, True)
    say(f"held-out loss (energy fraction span(P Z) misses): untrained {np.mean(held_out):.4f}, trained "
        f"{np.mean(held_trained):.4f}")
    say(f"PCG iterations over {eval_frames} frames: without the net {it_plain}, untrained net {it_untrained}, "
        f"trained net {it_trained}; warm-start residuals {['%.2e' % r for r in st['warm_relres']]}, "
        f"fallbacks {st['learned_fallbacks']}")
    trained.close()
    assert np.mean(held_trained) < 0.9 * np.mean(held_out)
    assert it_trained <= 1.05 * it_plain
    assert st["learned_fallbacks"] == 0


def test_non_finite_fields_fall_back_to_plain():
    from bd.solver import GpuSolver
    cfg = config_sheet(n=65, region=16, history=False)
    P = cfg["params"]
    P.lag_coarse = 0
    goals = np.array([g["goal"] for g in cfg["grips"]])
    ref = GpuSolver(cfg["scene"], P, cfg["grips"])
    _, it_ref = frames(ref, cfg, goals.copy(), 2)
    bad = GpuSolver(cfg["scene"], P, cfg["grips"])
    w = bad.learn_weights()
    w[:] = np.nan
    bad.set_learn_weights(w)
    bad.set_learned(True)
    _, it_bad = frames(bad, cfg, goals.copy(), 2)
    st = bad.stats()
    say(f"non-finite net: iterations {it_bad} (plain {it_ref}), fallbacks {st['learned_fallbacks']}")
    assert st["learned_fallbacks"] >= 3
    assert abs(sum(it_bad) - sum(it_ref)) <= 2

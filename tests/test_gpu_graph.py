"""Graph-captured substeps and batched independent solves on the GPU, through the C-ABI.

* a captured substep replayed frame after frame gives the eager substep's positions, velocities, PCG
  iteration counts and residuals bit for bit (the same kernels in the same order; the iterations
  queued past convergence return on the device's done flag; the capture's PCG budget is the eager
  solve's iteration limit, so an unconverged solve stops at the same iteration), on the 64 x 64 sheet
  (cfg 1) and on the folded sheet with self-contact and the floor; eager substeps are mixed between
  the replays on the same solver, and every replay is checked;
* a setter that changes what a substep launches makes the replay refuse until the next capture, and
  the capture refuses what needs the host between Newton iterations;
* a batch of uneven patches (each its own Newton / MG-PCG system) advanced by one graph launch per
  frame equals each patch's eager substeps bit for bit;
* the concurrent batch (each patch's substep on its own stream from a pool of host threads: eager, or
  the replay of a captured one) equals the patches' eager substeps run one after another bit for bit,
  with self-contact on some patches.
"""
import sys

import numpy as np
import pytest

from bd.configs import advance_grips, config1, config_fold, config_patches

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def same_frame(a, b, what):
    xa, xb = a.read(0), b.read(0)
    va, vb = a.read(1), b.read(1)
    assert np.isfinite(xa).all()
    assert np.array_equal(xa, xb), f"{what}: positions differ by {np.max(np.abs(xa - xb)):.3e}"
    assert np.array_equal(va, vb), f"{what}: velocities differ by {np.max(np.abs(va - vb)):.3e}"
    sa, sb = a.stats(), b.stats()
    n = sa["newton_iterations"]
    assert n == sb["newton_iterations"], what
    assert list(sa["pcg_iterations"][:n]) == list(sb["pcg_iterations"][:n]), (what, sa["pcg_iterations"], sb["pcg_iterations"])
    assert list(sa["rel_residual"][:n]) == list(sb["rel_residual"][:n]), what
    return sa


def run_pair(cfg, frames, floor=None):
    from bd.solver import GpuSolver
    S, P = cfg["scene"], cfg["params"]
    grips_a = [dict(g) for g in cfg.get("grips", [])]
    grips_b = [dict(g) for g in cfg.get("grips", [])]
    eager = GpuSolver(S, P, grips_a or None)
    graph = GpuSolver(S, P, grips_b or None)
    if floor is not None:
        eager.set_floor(floor)
        graph.set_floor(floor)
    vel = cfg.get("velocity", [])
    # a first eager frame on both (lazy set-up), then the capture
    for g in (eager, graph):
        g.encode_substep()
    same_frame(eager, graph, "first eager frame")
    graph.graph_capture(P.pcg_max_iterations)
    iters, worst = [], 0.0
    for k in range(frames):
        if grips_a:
            advance_grips(grips_a, vel)
            eager.set_grip_goals(np.array([g["goal"] for g in grips_a]))
            graph.set_grip_goals(np.array([g["goal"] for g in grips_a]))
        eager.encode_substep()
        if k % 3 == 2:
            graph.encode_substep()   # eager work between the replays
        else:
            graph.graph_substep()
        st = same_frame(eager, graph, f"frame {k}")
        iters.append(list(st["pcg_iterations"][:st["newton_iterations"]]))
        worst = max(worst, max(st["rel_residual"][:st["newton_iterations"]]))
    say(f"  {frames} frames, replays bit-exact against eager; PCG iterations per Newton {iters}, worst rel-res {worst:.2e}")
    return eager, graph


def test_graph_substep_matches_eager_on_the_sheet():
    say("gpu: cfg 1 sheet, eager and graph-captured substeps")
    eager, graph = run_pair(config1(), 9)
    eager.close()
    graph.close()


def test_graph_substep_matches_eager_with_contact_and_floor():
    cfg = config_fold()
    say("gpu: fold with self-contact and the floor, eager and graph-captured substeps")
    eager, graph = run_pair(cfg, 9, floor=cfg["floor_z"])
    eager.close()
    graph.close()


def test_graph_refusals_and_recapture():
    from bd.solver import GpuSolver
    cfg = config1()
    S, P = cfg["scene"], cfg["params"]
    g = GpuSolver(S, P, [dict(q) for q in cfg["grips"]])
    with pytest.raises(RuntimeError, match="no captured substep"):
        g.graph_substep()
    g.encode_substep()
    g.graph_capture(50)
    g.graph_substep()
    g.set_timestep(1 / 240)
    with pytest.raises(RuntimeError, match="configuration changed"):
        g.graph_substep()
    g.graph_capture(50)
    g.graph_substep()
    g.set_pcg(P.pcg_iterations, P.pcg_rel_tol, P.pcg_max_iterations)
    with pytest.raises(RuntimeError, match="configuration changed"):
        g.graph_substep()
    g.set_learned(1)
    with pytest.raises(RuntimeError, match="learned"):
        g.graph_capture(50)
    g.set_learned(0)
    g.close()
    P2 = cfg["params"]
    P2.newton_max = P2.newton_iterations + 2
    h = GpuSolver(S, P2, [dict(q) for q in cfg["grips"]])
    with pytest.raises(RuntimeError, match="optional Newton"):
        h.graph_capture(50)
    h.close()
    with pytest.raises(RuntimeError, match="positive"):
        GpuSolver(S, cfg["params"]).graph_capture(0)


def test_batch_of_uneven_patches_matches_eager():
    from bd.solver import Batch, GpuSolver
    cfg = config_patches(count=6, side_range=(32, 96), seed=11)
    sides = [p["side"] for p in cfg["patches"]]
    say(f"gpu: batch of {len(sides)} patches, sides {sides}")
    eager, batched, grips = [], [], []
    for p in cfg["patches"]:
        ga, gb = [dict(g) for g in p["grips"]], [dict(g) for g in p["grips"]]
        a, b = GpuSolver(p["scene"], p["params"], ga), GpuSolver(p["scene"], p["params"], gb)
        a.encode_substep()
        b.encode_substep()
        b.graph_capture(p["params"].pcg_max_iterations)
        eager.append(a)
        batched.append(b)
        grips.append(ga)
    batch = Batch(batched)
    for k in range(5):
        for p, a, b, ga in zip(cfg["patches"], eager, batched, grips):
            advance_grips(ga, p["velocity"])
            goals = np.array([g["goal"] for g in ga])
            a.set_grip_goals(goals)
            b.set_grip_goals(goals)
            a.encode_substep()
        batch.substep()
        its = []
        for i, (a, b) in enumerate(zip(eager, batched)):
            st = same_frame(a, b, f"frame {k} patch {i}")
            its.append(sum(st["pcg_iterations"][:st["newton_iterations"]]))
            assert max(st["rel_residual"][:st["newton_iterations"]]) <= 1.05e-6, (k, i, st)
        say(f"  frame {k}: PCG iterations per patch {its}, bit-exact")
    # a solver captured again no longer matches the batch's copy of its graph
    batched[0].graph_capture(cfg["patches"][0]["params"].pcg_max_iterations)
    with pytest.raises(RuntimeError, match="captured again or changed"):
        batch.substep()
    batch.close()
    for g in eager + batched:
        g.close()


def test_concurrent_batch_matches_eager_substeps():
    from bd.solver import Batch, GpuSolver
    cfg = config_patches(count=7, side_range=(32, 120), seed=5)
    sides = [p["side"] for p in cfg["patches"]]
    say(f"gpu: concurrent batch of {len(sides)} patches, sides {sides}, 3 host threads")
    eager, batched, grips = [], [], []
    for i, p in enumerate(cfg["patches"]):
        params = p["params"]
        params.self_contact = i % 2 == 1
        ga, gb = [dict(g) for g in p["grips"]], [dict(g) for g in p["grips"]]
        eager.append(GpuSolver(p["scene"], params, ga))
        batched.append(GpuSolver(p["scene"], params, gb))
        grips.append((ga, gb))
    # every third patch replays a captured substep (budget = the eager limit, so bit-exact)
    for i, b in enumerate(batched):
        if i % 3 == 2:
            b.encode_substep()
            b.graph_capture(cfg["patches"][i]["params"].pcg_max_iterations)
    for i, a in enumerate(eager):
        if i % 3 == 2:
            a.encode_substep()
    batch = Batch(batched, threads=3)
    for k in range(4):
        for p, a, (ga, gb) in zip(cfg["patches"], eager, grips):
            advance_grips(ga, p["velocity"])
            advance_grips(gb, p["velocity"])
        for a, b, (ga, gb) in zip(eager, batched, grips):
            a.set_grip_goals(np.array([g["goal"] for g in ga]))
            b.set_grip_goals(np.array([g["goal"] for g in gb]))
            a.encode_substep()
        batch.substep()
        batch.sync()
        its = []
        for i, (a, b) in enumerate(zip(eager, batched)):
            st = same_frame(a, b, f"frame {k} patch {i}")
            its.append(sum(st["pcg_iterations"][:st["newton_iterations"]]))
            assert max(st["rel_residual"][:st["newton_iterations"]]) <= 1.05e-6, (k, i, st)
        say(f"  frame {k}: PCG iterations per patch {its}, bit-exact")
    batch.close()
    for g in eager + batched:
        g.close()

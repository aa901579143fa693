"""The 1024 x 1024 workload with self-contact on the GPU, stage by stage, through the C-ABI.

The sheet has 1,048,576 nodes, one more than the reference's 20-bit reciprocal key can name
(src/simulation/gpu/selfcontact.ts:84); here the key is 64 bits (share code over a 32-bit node). After
two frames of the mixed method (the hierarchy rebuilt every Newton iteration, the 17 x 17 coarsest level
solved exactly in FP64), the third frame is checked on the GPU's own state:

* detection on the predicted step: cell starts and counts, every cell's records, the K candidate
  faces, reciprocal keys, row marks, own and reciprocal bounds, bit for bit against the FP32
  restatement (oracle/contact_oracle.c);
* the contact terms after projectStart: ext and the trust-region bounds, bit for bit;
* the first Newton iteration's fine operator and rhs (with those terms, the active bounds and the two
  grips) against the FP64 oracle at the reference's bars (5e-5 and 1e-3, tests/gpu/correctness.ts:
  180-196), the Galerkin levels built from the GPU's fine operator at 1e-5;
* the mixed method's PCG to rel-res 1e-6: the true residual of the GPU's FP32 iterate, recomputed in
  FP64 with the same operator, is at most 1.05e-6.
"""
import copy
import sys

import numpy as np
import pytest

import oracle as O
from bd.capi import BD_BUF_DELTA, BD_BUF_DX, BD_BUF_RHS, BD_COARSE_EXACT, BD_PRECISION_FP32
from bd.configs import config_sheet

from test_gpu_contact import f32_gnorm
from test_gpu_parity import rel_to_scale

pytestmark = pytest.mark.gpu
K = 4
T2 = 1 << 17


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def records_in_cell_order(cells, items):
    """The records of every occupied cell, sorted by (cell, face id): within a cell the fill order is
    that of the atomics. The starts must be the exclusive scan of the counts in cell order."""
    starts, counts = cells[:T2].astype(np.int64), cells[T2:].astype(np.int64)
    scan = np.concatenate([[0], np.cumsum(counts)[:-1]])
    assert np.array_equal(starts, scan), "cell starts are the exclusive scan of the counts"
    total = int(counts.sum())
    cell_of = np.repeat(np.arange(T2), counts)
    r = np.asarray(items[:total])
    face = r[:, 7].view(np.uint32)
    order = np.lexsort((face, cell_of))
    return r[order].view(np.uint32)


def test_sheet1024_contact_and_stages():
    from bd.solver import GpuSolver
    cfg = config_sheet(1024, self_contact=True)
    S = cfg["scene"]
    P = copy.copy(cfg["params"])
    P.coarse_solve = BD_COARSE_EXACT
    P.max_coarsest_nodes = 400
    P.lag_coarse = 0
    P.precision = BD_PRECISION_FP32
    grips = [dict(g) for g in cfg["grips"]]
    assert S.node_count == 1 << 20
    say(f"gpu: {S.node_count}-node sheet with self-contact, mixed method, 2 frames")
    gpu = GpuSolver(S, P, grips)
    goals = np.array([g["goal"] for g in grips], dtype=np.float64)
    vel = np.array(cfg["velocity"], dtype=np.float64)
    for k in range(2):
        goals = goals + vel
        gpu.set_grip_goals(goals)
        gpu.encode_substep()
        st = gpu.stats()
        say(f"  frame {k + 1}: pcg {st['pcg_iterations'][:st['newton_iterations']]}, "
            f"rel-res {max(st['rel_residual'][:st['newton_iterations']]):.2e}")
        assert max(st["rel_residual"][:st["newton_iterations"]]) <= 1.05e-6

    # frame 3, stage by stage
    goals = goals + vel
    gpu.set_grip_goals(goals)
    x = gpu.x.astype(np.float32)
    v = gpu.read(1)
    flags = gpu.read(15)
    say("gpu: predict + contact detection on frame 3")
    gpu.encode_predict()
    dx = gpu.vec(BD_BUF_DX).astype(np.float32)
    gpu.encode_contact_detect()
    T = O.ContactTwin(S, x, dx, flags=flags)
    cells, items, pairs, cmeta = T.detect()
    gcells = gpu.cells()
    assert np.array_equal(gcells, cells), "cell starts / counts"
    assert np.array_equal(records_in_cell_order(gcells, gpu.items()), records_in_cell_order(cells, items)), "records"
    assert np.array_equal(gpu.pairs(), pairs), "candidate faces, reciprocal keys, row marks"
    assert np.array_equal(gpu.cmeta().view(np.uint32), cmeta.view(np.uint32)), "own and reciprocal bounds"
    n_pairs = int((pairs[:, :K] != 0xFFFFFFFF).sum())
    say(f"detect: {int(cells[T2:].sum())} face records, {n_pairs} candidate faces, "
        f"{int((cmeta[:, 0, 3] < 1e8).sum())} bounded nodes, {int((cmeta[:, 1, 3] < 1e8).sum())} reciprocal")
    assert n_pairs > 0

    say("gpu: projectStart + contact terms + assembly with hierarchy")
    gpu.encode_project_start()
    dxs = gpu.vec(BD_BUF_DX).astype(np.float32)
    gpu.encode_contact_terms()
    cm = gpu.cmeta().copy()
    # the terms on the same projected step and the bounds that projectStart left
    T = O.ContactTwin(S, x, dxs, flags=flags)
    cm_twin = cm.copy()
    _, ext = T.terms(pairs, cm_twin, np.float32(P.h), f32_gnorm(P.gravity))
    assert np.array_equal(gpu.ext().view(np.uint32), ext.view(np.uint32)), "ext"
    assert np.array_equal(gpu.cmeta().view(np.uint32), cm_twin.view(np.uint32)), "trust-region bounds"
    gpu.encode_assemble(True)

    cpu = O.Oracle(S, P, grips)
    for g, goal in zip(grips, goals):
        g["goal"] = goal
    cpu.set_grips(grips)
    cpu.x[:] = x.astype(np.float64).ravel()
    cpu.v[:] = v
    cpu.rest[:] = gpu.read(12)
    cpu.intact[:] = gpu.read(13)
    cpu.cut_mask[:] = gpu.read(14)
    cpu.seg_flags[:] = flags
    cpu.fail_site[:] = gpu.read(16)
    cpu.tails[:] = gpu.read(17)
    cpu.damage[:] = gpu.read(18)
    cpu.predict()
    cpu.dx[:] = dxs.astype(np.float64).ravel()
    cpu.set_ext(gpu.ext().astype(np.float64))
    cpu.set_cmeta(gpu.cmeta().astype(np.float64))
    cpu.set_floor(None)
    cpu.assemble()
    Ag, Ac = gpu.stencil(0), cpu.stencil(0)
    e = rel_to_scale(Ag.reshape(len(Ag), -1), Ac.reshape(len(Ac), -1))
    say(f"fine operator maxAbs/scale = {e:.2e}")
    assert e < 5e-5
    b = gpu.read(BD_BUF_RHS)
    e = rel_to_scale(b, cpu.rhs())
    say(f"rhs maxAbs/scale = {e:.2e}")
    assert e < 1e-3
    fixed0 = gpu.fixed(0)
    assert np.array_equal(fixed0, cpu.fixed(0))
    cpu.load_fine_operator(Ag, fixed0)
    levels = gpu.stats()["level_nodes"]
    assert cpu.level_count() == len(levels)
    for k in range(1, cpu.level_count()):
        Ak, Ck = gpu.stencil(k), cpu.stencil(k)
        e = rel_to_scale(Ak.reshape(len(Ak), -1), Ck.reshape(len(Ck), -1))
        say(f"galerkin level {k} ({levels[k]} nodes) maxAbs/scale = {e:.2e}")
        assert e < 1e-5

    say("gpu: PCG to rel-res 1e-6")
    gpu.encode_pcg()
    st = gpu.stats()
    d = gpu.read(BD_BUF_DELTA)
    fixed = np.repeat(fixed0 == 1, 3)
    b[fixed] = 0
    r = b - cpu.apply(0, d)
    r[fixed] = 0
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    say(f"pcg iterations {st['pcg_iterations'][0]}, reported rel-res {st['rel_residual'][0]:.2e}, FP64 recomputed {rel:.2e}")
    assert st["rel_residual"][0] <= 1e-6
    assert rel < 1.05e-6
    gpu.close()

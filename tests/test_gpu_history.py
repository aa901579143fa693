"""Bit-exact history updates: yarn damage codes, rupture events, pull-outs and crossing slip on the
GPU equal the FP64 oracle's on the same committed positions (yarnField.ts / solver.ts:659-740)."""
import sys

import numpy as np
import pytest

import oracle as O
from bd.capi import make_params
from bd.scene import compile_scene, SEG_BROKEN

from test_gpu_parity import corr_coupon, say

pytestmark = pytest.mark.gpu


def stretched(S, amp_u, amp_v, seed):
    """Positions with a graded, nonuniform stretch along both yarn families (FP32-rounded)."""
    L = S.layers[0]
    rng = np.random.default_rng(seed)
    tu = np.array([L.mapA[0], L.mapA[2], 0.0])
    tv = np.array([L.mapA[1], L.mapA[3], 0.0])
    tu /= np.linalg.norm(tu)
    tv /= np.linalg.norm(tv)
    eu = amp_u * rng.random(L.nx - 1)
    ev = amp_v * rng.random(L.ny - 1)
    gu = np.concatenate([[0], np.cumsum(eu * L.material["pitchU"])])
    gv = np.concatenate([[0], np.cumsum(ev * L.material["pitchV"])])
    j, i = np.meshgrid(np.arange(L.ny), np.arange(L.nx), indexing="ij")
    x = S.position.copy()
    x += gu[i.ravel()][:, None] * tu + gv[j.ravel()][:, None] * tv
    return x.astype(np.float32).astype(np.float64)


def compare_state(gpu, cpu, label):
    assert np.array_equal(gpu.read(18), cpu.damage), f"{label}: damage codes"
    assert np.array_equal(gpu.read(14), cpu.cut_mask), f"{label}: cut masks"
    assert np.array_equal(gpu.read(15), cpu.seg_flags), f"{label}: span flags"
    assert np.array_equal(gpu.read(13), cpu.intact), f"{label}: intact fractions"
    assert np.array_equal(gpu.read(16), cpu.fail_site), f"{label}: fail sites"
    assert np.array_equal(gpu.read(17), cpu.tails), f"{label}: tails"
    assert np.array_equal(gpu.read(12), cpu.rest), f"{label}: rest lengths"


def test_damage_field_and_slip_bit_exact():
    from bd.solver import GpuSolver
    S = compile_scene([corr_coupon()])
    P = make_params()
    cpu = O.Oracle(S, P)
    gpu = GpuSolver(S, P)
    total_events = 0
    for round_, (au, av) in enumerate(((0.1, 0.3), (0.3, 0.6), (0.6, 1.0))):
        x = stretched(S, au, av, 100 + round_)
        cpu.x[:] = x.ravel()
        gpu.set_state(x)
        say(f"gpu: damage round {round_} (stretch {au}, {av})")
        cpu.update_damage()
        gpu.encode_damage()
        compare_state(gpu, cpu, f"damage round {round_}")
        say(f"gpu: slip round {round_}")
        cpu.update_slip()
        gpu.encode_slip()
        compare_state(gpu, cpu, f"slip round {round_}")
        st = gpu.stats()
        assert st["ruptures"] == cpu.broken_count
        assert st["pulled_out"] == cpu.pulled_out_count
        total_events = cpu.broken_count + int(np.count_nonzero(cpu.damage >= 254))
    broken = int(np.count_nonzero(cpu.seg_flags & SEG_BROKEN))
    say(f"ruptures {cpu.broken_count}, broken spans {broken}, separated crossings {int(np.count_nonzero(cpu.damage >= 254))}, "
        f"slipped out {cpu.pulled_out_count}")
    # the scenario must actually exercise the events it claims to check
    assert cpu.broken_count > 0 and total_events > cpu.broken_count and cpu.pulled_out_count > 0
    assert int(np.count_nonzero((cpu.damage > 0) & (cpu.damage < 254))) > 0

"""The solver's deterministic FP64 grid reductions (one partial per block, summed in block order by
the last block to finish, with no fence between the blocks: each partial is an agent-scope atomic
store whose completion precedes the block's ticket): x^T A x of bd_apply_noalloc on a 513 x 513 sheet's
fine operator, for two vectors in alternation over 600 calls, so that a partial read before its block
had stored it would leave the other vector's value in the sum. Every result must equal the first one
for its vector bit for bit, and match the FP64 product of x with the returned y."""
import sys

import numpy as np
import pytest

from test_zz_no_alloc import DeviceArray, hip_runtime

pytestmark = pytest.mark.gpu


def test_grid_reduction_repeats_bit_for_bit():
    from bd.configs import config_sheet
    from bd.solver import GpuSolver
    cfg = config_sheet(n=513, region=64, history=False)
    gpu = GpuSolver(cfg["scene"], cfg["params"], cfg["grips"])
    hip = hip_runtime()
    gpu.encode_predict()
    gpu.encode_assemble(True)
    n = gpu.stats()["level_nodes"][0]
    Np = (n + 63) // 64 * 64
    rng = np.random.default_rng(11)
    xs = [DeviceArray(hip, rng.standard_normal(3 * Np).astype(np.float32)) for _ in range(2)]
    y = DeviceArray(hip, np.zeros(3 * Np, dtype=np.float32))
    d = DeviceArray(hip, np.zeros(1, dtype=np.float64))
    first = [None, None]
    for it in range(600):
        k = it % 2
        rc = gpu.lib.bd_apply_noalloc(gpu.h, 0, xs[k].ptr.value, y.ptr.value, d.ptr.value, None)
        assert rc == 0, gpu.lib.bd_last_error(gpu.h).decode()
        v = float(d.read()[0])
        if first[k] is None:
            first[k] = v
            yk = y.read().reshape(3, Np)[:, :n].astype(np.float64)
            xk = xs[k].host.reshape(3, Np)[:, :n].astype(np.float64)
            ref = float(np.sum(xk * yk))
            print(f"vector {k}: x^T A x = {v:.12e}, FP64 x.y = {ref:.12e}", flush=True)
            sys.stdout.flush()
            assert abs(v - ref) <= 1e-6 * abs(ref)
        assert v == first[k], f"call {it}: {v!r} against {first[k]!r}"
    assert first[0] != first[1]
    for a in xs + [y, d]:
        a.free()
    gpu.close()

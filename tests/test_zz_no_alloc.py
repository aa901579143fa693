"""bd_apply_noalloc (no allocation; scratch in a fixed device array) against bd_apply (a workspace
allocated per call): the same level operator applied to the same device vectors gives the same y
and x^T A x bit for bit, on two calls in a row, the second on a smaller level."""
import sys

import numpy as np
import pytest

from bd.capi import BD_BUF_LEVEL_X

pytestmark = pytest.mark.gpu


def test_apply_noalloc_matches_apply():
    import torch
    from bd.configs import config_sheet
    from bd.solver import GpuSolver
    cfg = config_sheet(n=65, region=16, history=False)
    gpu = GpuSolver(cfg["scene"], cfg["params"], cfg["grips"])
    gpu.encode_predict()
    gpu.encode_assemble(True)
    levels = gpu.stats()["level_nodes"]
    assert len(levels) >= 2 and levels[1] < levels[0]
    gen = torch.Generator(device="cuda").manual_seed(3)
    for level in (0, 1):  # the second call on the smaller level
        Np = (levels[level] + 63) // 64 * 64
        x = torch.randn(3 * Np, device="cuda", dtype=torch.float32, generator=gen)
        outs = []
        for fn in (gpu.lib.bd_apply, gpu.lib.bd_apply_noalloc):
            y = torch.full((3 * Np,), float("nan"), device="cuda", dtype=torch.float32)
            d = torch.zeros(1, device="cuda", dtype=torch.float64)
            torch.cuda.synchronize()
            rc = fn(gpu.h, level, x.data_ptr(), y.data_ptr(), d.data_ptr(), None)
            assert rc == 0, gpu.lib.bd_last_error(gpu.h).decode()
            torch.cuda.synchronize()
            outs.append((y.cpu().numpy(), float(d.cpu().numpy()[0])))
        (y1, d1), (y2, d2) = outs
        print(f"level {level} ({levels[level]} nodes): x^T A x = {d1:.9e} (apply) {d2:.9e} (noalloc)", flush=True)
        sys.stdout.flush()
        n = levels[level]
        assert np.isfinite(y1[:n]).all() and d1 != 0.0
        assert np.array_equal(y1.reshape(3, Np)[:, :n], y2.reshape(3, Np)[:, :n])
        assert d1 == d2
    gpu.close()

"""bd_apply_noalloc (no allocation; scratch in a fixed device array) against bd_apply (a workspace
allocated per call): the same level operator applied to the same device vectors gives the same y
and x^T A x bit for bit, on two calls in a row, the second on a smaller level.

The device vectors come from the HIP runtime that libbd.so itself links (hipMalloc / hipMemcpy through
ctypes), so the process holds one HIP runtime whatever ran before it."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip_runtime():
    """The HIP runtime already loaded with libbd.so (its soname), with the few entry points used here."""
    hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    for f in (hip.hipMalloc, hip.hipFree, hip.hipMemcpy, hip.hipDeviceSynchronize):
        f.restype = C.c_int
    return hip


class DeviceArray:
    def __init__(self, hip, host):
        self.hip, self.host = hip, np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), self.host.nbytes) == 0
        assert hip.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def read(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        assert self.hip.hipFree(self.ptr) == 0


def test_apply_noalloc_matches_apply():
    from bd.configs import config_sheet
    from bd.solver import GpuSolver
    cfg = config_sheet(n=65, region=16, history=False)
    gpu = GpuSolver(cfg["scene"], cfg["params"], cfg["grips"])
    hip = hip_runtime()
    gpu.encode_predict()
    gpu.encode_assemble(True)
    levels = gpu.stats()["level_nodes"]
    assert len(levels) >= 2 and levels[1] < levels[0]
    rng = np.random.default_rng(3)
    for level in (0, 1):  # the second call on the smaller level
        Np = (levels[level] + 63) // 64 * 64
        x = DeviceArray(hip, rng.standard_normal(3 * Np).astype(np.float32))
        outs = []
        for fn in (gpu.lib.bd_apply, gpu.lib.bd_apply_noalloc):
            y = DeviceArray(hip, np.full(3 * Np, np.nan, dtype=np.float32))
            d = DeviceArray(hip, np.zeros(1, dtype=np.float64))
            assert hip.hipDeviceSynchronize() == 0
            rc = fn(gpu.h, level, x.ptr.value, y.ptr.value, d.ptr.value, None)
            assert rc == 0, gpu.lib.bd_last_error(gpu.h).decode()
            assert hip.hipDeviceSynchronize() == 0
            outs.append((y.read(), float(d.read()[0])))
            y.free()
            d.free()
        x.free()
        (y1, d1), (y2, d2) = outs
        print(f"level {level} ({levels[level]} nodes): x^T A x = {d1:.9e} (apply) {d2:.9e} (noalloc)", flush=True)
        sys.stdout.flush()
        n = levels[level]
        assert np.isfinite(y1[:n]).all() and d1 != 0.0
        assert np.array_equal(y1.reshape(3, Np)[:, :n], y2.reshape(3, Np)[:, :n])
        assert d1 == d2
    gpu.close()

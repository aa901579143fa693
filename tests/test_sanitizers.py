"""The CPU code under AddressSanitizer and UndefinedBehaviorSanitizer, and run-twice determinism.

* The FP64 oracle (oracle/bd_oracle.c, oracle/contact_oracle.c) and the host topology search
  (break-that-denim_amd/csrc/topology.cpp) are built with -fsanitize=address,undefined into a
  temporary directory and driven in a child process (the sanitizer runtimes preloaded): substeps of
  the worn bundled coupon with the damage field and slip, a pull to rupture, contact detection and
  terms on the pushed fold with the oracle's contact substep, and the topology search on random span
  graphs. Any sanitizer report fails the test.
* The oracle run twice on the same inputs gives the same state bit for bit (solver and contact).
"""
import os
import shutil
import subprocess
import sys
import textwrap
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from bd.configs import config_fold
from bd.capi import make_params
from bd.scene import compile_scene, coupon

ROOT = Path(__file__).resolve().parent.parent

CHILD = textwrap.dedent('''
    import ctypes as C, sys
    import numpy as np
    sys.path.insert(0, "{root}/break-that-denim_amd"); sys.path.insert(0, "{root}/oracle")
    import oracle as O
    from bd.capi import make_params
    from bd.configs import config_fold
    from bd.scene import compile_scene, coupon

    # the worn bundled coupon with the damage field and slip
    S = compile_scene([coupon(24, 20, bundle=[3, 2], angle=0.3, compressU=0.03, compressV=0.02, shear=0.02,
                              collarRows=2, seed=7, perturbation=0.4e-3)])
    P = make_params(newton_iterations=2, pcg_iterations=3, history=True)
    o = O.Oracle(S, P)
    for _ in range(4):
        o.substep()
    assert np.isfinite(o.x).all()
    # a pull to rupture: the free bottom rows dragged down
    y = o.x.reshape(-1, 3)
    low = y[:, 2] < y[:, 2].min() + 2e-3
    for _ in range(6):
        o.x.reshape(-1, 3)[low, 2] -= 2e-3
        o.substep()
    print("coupon: broken", o.broken_count, flush=True)

    # contact: detection and terms on the pushed fold, then the oracle's contact substep
    cfg = config_fold(nx=16, ny=24)
    Sf, Pf = cfg["scene"], cfg["params"]
    x = Sf.position.astype(np.float32)
    dx = np.zeros_like(x)
    dx[x[:, 2] > 1.5e-3, 2] = -2.5e-3
    T = O.ContactTwin(Sf, x, dx)
    cells, items, pairs, cmeta = T.detect()
    cm = cmeta.copy()
    acc, ext = T.terms(pairs, cm, np.float32(Pf.h), np.float32(9.81))
    print("fold: candidates", int((pairs[:, :4] != 0xFFFFFFFF).sum()), flush=True)
    of = O.Oracle(Sf, Pf)
    of.predict()
    x32 = of.x.reshape(-1, 3).astype(np.float32)
    _, _, pr, cmt = O.ContactTwin(Sf, x32, of.dx.reshape(-1, 3).astype(np.float32)).detect()
    of.set_cmeta(cmt.astype(np.float64))
    of.set_floor(cfg["floor_z"])
    of.project_start()
    for it in range(2):
        _, e = O.ContactTwin(Sf, x32, of.dx.reshape(-1, 3).astype(np.float32)).terms(pr, cmt, np.float32(Pf.h), np.float32(9.81))
        of.set_ext(e.astype(np.float64))
        of.set_cmeta(cmt.astype(np.float64))
        of.newton_step(it == 0)
    of.finalize()
    assert np.isfinite(of.x).all()

    # the topology search on random span graphs of two layers
    topo = C.CDLL("{topo}")
    i32, u8, u32 = C.POINTER(C.c_int32), C.POINTER(C.c_uint8), C.POINTER(C.c_uint32)
    rng = np.random.default_rng(5)
    for trial in range(4):
        nx = np.array([30, 12], np.int32); ny = np.array([20, 9], np.int32); off = np.array([0, 600], np.int32)
        N = 708
        bits = (rng.random(N) > 0.2).astype(np.uint8) | ((rng.random(N) > 0.2).astype(np.uint8) << 1)
        bits[rng.random(N) < 0.05] |= 4
        mounted = np.zeros(N, np.uint8); mounted[:30] = 1
        anchors = rng.integers(0, N, 5).astype(np.int32)
        freed = np.zeros(N, np.uint8); pof = np.zeros(N, np.uint32)
        start = np.zeros(N + 1, np.uint32); nodes = np.zeros(N + 1, np.uint32); count = C.c_int32()
        rc = topo.bd_topology_search(N, 2, nx.ctypes.data_as(i32), ny.ctypes.data_as(i32), off.ctypes.data_as(i32),
                                     bits.ctypes.data_as(u8), mounted.ctypes.data_as(u8), anchors.ctypes.data_as(i32), 5,
                                     freed.ctypes.data_as(u8), pof.ctypes.data_as(u32), start.ctypes.data_as(u32),
                                     nodes.ctypes.data_as(u32), C.byref(count))
        assert rc >= 0
    print("topology: ok", flush=True)
''')


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None, reason="needs gcc and g++")
def test_oracle_and_host_code_under_sanitizers(tmp_path):
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    ubsan = subprocess.run(["gcc", "-print-file-name=libubsan.so"], capture_output=True, text=True).stdout.strip()
    if not (os.path.isabs(asan) and os.path.isabs(ubsan)):
        pytest.skip("the sanitizer runtimes are not installed")
    san = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fPIC",
           "-shared"]
    lib = tmp_path / "liboracle_san.so"
    subprocess.run(["gcc", *san, "-std=c11", "-ffp-contract=off", "-o", str(lib), str(ROOT / "oracle" / "bd_oracle.c"),
                    str(ROOT / "oracle" / "contact_oracle.c"), "-lm"], check=True)
    topo = tmp_path / "libtopo_san.so"
    subprocess.run(["g++", *san, "-std=c++17", "-I", str(ROOT / "include"), "-o", str(topo),
                    str(ROOT / "break-that-denim_amd" / "csrc" / "topology.cpp")], check=True)
    script = tmp_path / "child.py"
    script.write_text(CHILD.format(root=ROOT, topo=topo))
    env = dict(os.environ, BD_ORACLE_LIB=str(lib), LD_PRELOAD=f"{asan}:{ubsan}", PYTHONMALLOC="malloc",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, str(script)], env=env, capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    print(out[-3000:])
    assert r.returncode == 0, out[-3000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out
    assert "topology: ok" in out


def test_oracle_runs_twice_bit_identical():
    S = compile_scene([coupon(24, 20, bundle=[3, 2], seed=7, perturbation=0.4e-3)])
    P = make_params(newton_iterations=2, pcg_iterations=3, history=True)
    runs = []
    for _ in range(2):
        o = O.Oracle(S, P)
        for _ in range(5):
            o.substep()
        runs.append((o.x.copy(), o.v.copy(), o.damage.copy(), o.rest.copy()))
    for a, b in zip(*runs):
        assert np.array_equal(a, b)
    cfg = config_fold(nx=16, ny=24)
    x = cfg["scene"].position.astype(np.float32)
    dx = np.zeros_like(x)
    dx[x[:, 2] > 1.5e-3, 2] = -2.5e-3
    first = O.ContactTwin(cfg["scene"], x, dx).detect()
    second = O.ContactTwin(cfg["scene"], x, dx).detect()
    for a, b in zip(first, second):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))

"""The Node host on the GPU: the same scene, parameters and grip goals through the N-API addon
(tests/node/run_substeps.js, GpuSolver of break-that-denim_amd/node/gpu_solver.js) and through the
Python host give the same positions and velocities bit for bit (both call the same C-ABI)."""
import json
import subprocess
import sys

import numpy as np
import pytest

import node_addon
from bd.capi import make_params
from bd.configs import corner_grip, sheet_scene
from bd.material import linear_sheet_material

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not node_addon.toolchain(), reason="node not installed")]


def test_node_host_matches_python_host(tmp_path):
    from bd.solver import GpuSolver
    if not node_addon.ADDON.exists():
        pytest.skip("the addon is built by __graft_entry__.build() (node-gyp)")
    pitch = 0.5 / 31
    mat = linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / 2.4)
    S = sheet_scene(32, pitch, mat)
    grip = corner_grip(S, 1, 0, k=2e4)
    P = make_params(h=1 / 120, newton_iterations=3, pcg_rel_tol=1e-6, pcg_max_iterations=100, history=True)
    goals = [np.asarray(grip["goal"]) + np.array([2e-3 * (k + 1), 0, 0]) for k in range(4)]
    node_addon.export(tmp_path / "in.json", S, P, [grip], [g[None] for g in goals], steps=4)
    r = subprocess.run(["node", str(node_addon.ROOT / "tests" / "node" / "run_substeps.js"), str(tmp_path / "in.json"),
                        str(tmp_path / "out.json")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    sys.stdout.flush()
    assert r.returncode == 0, r.stderr
    out = json.loads((tmp_path / "out.json").read_text())
    gpu = GpuSolver(S, P, [grip])
    for g in goals:
        gpu.set_grip_goals(g[None])
        gpu.encode_substep()
    xn, vn = np.array(out["x"]), np.array(out["v"])
    assert np.array_equal(xn, gpu.read(0)) and np.array_equal(vn, gpu.read(1))
    assert np.max(np.abs(xn - S.position.ravel())) > 1e-3  # the grip moved the sheet
    gpu.close()

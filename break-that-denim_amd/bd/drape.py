"""Settled-drape state format and the drape settle (src/simulation/gpu/settle.ts:1-62).

File format (little-endian): four u32 [magic 0x50415244, version 1, node count, sceneHash], then
4 f32 per node (x, y, z and the reference's state lane w: 0 cloth, > 0.5 detached, > 1.5 retired).
A drape is accepted only for the scene it was computed from (same node count and sceneHash).
"""
from __future__ import annotations

import numpy as np

from .scene import scene_hash

DRAPE_MAGIC = 0x50415244
DRAPE_VERSION = 1
# settleDrape (settle.ts:24-25): converged, damage-free substeps, kinetic damping every 10 substeps
SETTLE_STEPS = 420
SETTLE_DAMPED = 360


def encode_drape(scene, positions) -> bytes:
    """encodeDrape (settle.ts:8-13): header + 4 f32 per node (w = 0 where only xyz is given)."""
    p = np.asarray(positions, dtype=np.float32).reshape(scene.node_count, -1)
    if p.shape[1] == 3:
        p = np.concatenate([p, np.zeros((scene.node_count, 1), np.float32)], axis=1)
    header = np.array([DRAPE_MAGIC, DRAPE_VERSION, scene.node_count, scene_hash(scene)], dtype=np.uint32)
    return header.tobytes() + np.ascontiguousarray(p, dtype=np.float32).tobytes()


def decode_drape(scene, data: bytes):
    """The (N, 4) f32 positions of a drape computed for this scene, or None (settle.ts:16-22)."""
    if len(data) != 16 + 16 * scene.node_count:
        return None
    h = np.frombuffer(data[:16], dtype=np.uint32)
    if h[0] != DRAPE_MAGIC or h[1] != DRAPE_VERSION or h[2] != scene.node_count or h[3] != scene_hash(scene):
        return None
    return np.frombuffer(data[16:], dtype=np.float32).reshape(scene.node_count, 4).copy()


def node_states(w):
    """The node states a drape's w lane holds: 0 cloth, 1 detached (w > 0.5), 2 retired (w > 1.5)."""
    w = np.asarray(w)
    return np.where(w > 1.5, 2, np.where(w > 0.5, 1, 0)).astype(np.uint8)


def load_drape(solver, data: bytes) -> bool:
    """loadDrape (settle.ts:16-22): upload a drape that matches the solver's scene; False otherwise.
    Detached and retired nodes keep their state (the w lane); detached pieces are then labelled
    again from the span graph (Topology.request)."""
    p = decode_drape(solver.scene, data)
    if p is None:
        return False
    solver.set_state(p[:, :3].astype(np.float64))
    st = node_states(p[:, 3])
    if st.any():
        solver.set_node_state(st)
        solver.update_topology()
    return True


def settle_drape(solver, steps=SETTLE_STEPS, damped=SETTLE_DAMPED, newton=5, pcg_iterations=2):
    """settleDrape (settle.ts:27-62) on a GpuSolver: `steps` substeps of `newton` Newton x
    `pcg_iterations` PCG iterations with the history updates off, the velocity cleared every 10
    substeps before `damped`, and cleared once more at the end. The solver's Newton and PCG
    schedule and history flag are restored afterwards."""
    p = solver.params
    saved = (p.newton_iterations, p.newton_max, p.pcg_iterations, p.pcg_rel_tol, p.pcg_max_iterations)
    solver.set_pcg(pcg_iterations, 0.0, p.pcg_max_iterations)
    solver.set_history_enabled(False)
    zero_v = np.zeros(3 * solver.N)
    try:
        for k in range(steps):
            solver.encode_substep(newton, newton)
            if k % 10 == 9 and k < damped:
                solver.write_velocity(zero_v)
    finally:
        solver.set_pcg(saved[2], saved[3], saved[4])
        solver.set_history_enabled(bool(p.history))
    solver.write_velocity(zero_v)
    solver.sync()

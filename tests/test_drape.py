"""Settled-drape format and state (src/simulation/gpu/settle.ts), on the reference's own settled
drape of its default scene (tests/golden/drape_default.bin, the product's public/drape.bin)."""
from pathlib import Path

import numpy as np

import oracle as O
from bd.capi import make_params
from bd.drape import DRAPE_MAGIC, decode_drape, encode_drape
from bd.scene import compile_scene, default_scene_specs

GOLDEN = Path(__file__).resolve().parent / "golden" / "drape_default.bin"


def test_drape_header_matches_the_compiled_default_scene():
    S = compile_scene(default_scene_specs())
    data = GOLDEN.read_bytes()
    h = np.frombuffer(data[:16], dtype=np.uint32)
    assert list(h) == [DRAPE_MAGIC, 1, 20930, 4202117243]
    p = decode_drape(S, data)
    assert p is not None and p.shape == (20930, 4)
    assert np.isfinite(p).all() and (p[:, 3] == 0).all()
    # the settle moved the cloth by millimetres, pinned nodes not at all
    d = np.linalg.norm(p[:, :3] - S.position, axis=1)
    assert 1e-3 < d.max() < 2e-2
    assert np.max(d[S.pinned == 1]) < 1e-6


def test_encode_decode_round_trip_and_rejection():
    S = compile_scene(default_scene_specs())
    p = decode_drape(S, GOLDEN.read_bytes())
    data = encode_drape(S, p)
    assert data == GOLDEN.read_bytes()
    assert decode_drape(S, data[:-16]) is None
    bad = bytearray(data)
    bad[12] ^= 1                                  # another sceneHash
    assert decode_drape(S, bytes(bad)) is None


def test_settled_drape_is_near_equilibrium_of_the_oracle():
    """At the reference's settled state (zero velocity) the oracle's Newton rhs, the net force on
    each free node, is two orders of magnitude below its value at the compiled start state."""
    S = compile_scene(default_scene_specs())
    p = decode_drape(S, GOLDEN.read_bytes())
    params = make_params(history=False)
    free = S.pinned == 0
    weight = S.mass[free] * 9.81

    def residual(x):
        o = O.Oracle(S, params)
        o.x[:] = np.asarray(x, dtype=np.float64).ravel()
        o.v[:] = 0
        o.predict()
        return np.linalg.norm(o.rhs().reshape(-1, 3)[free], axis=1) / weight

    settled, start = residual(p[:, :3]), residual(S.position)
    assert np.median(settled) < 1e-2 * np.median(start)
    assert np.percentile(settled, 90) < 1e-2 * np.percentile(start, 90)


def test_node_states_from_the_w_lane():
    """The drape's w lane holds the node state (state.wgsl): 0 cloth, > 0.5 detached, > 1.5 retired."""
    from bd.drape import node_states
    assert node_states([0.0, 0.4, 0.6, 1.0, 1.5, 1.6, 2.0]).tolist() == [0, 0, 1, 1, 1, 2, 2]

"""Host scene compiler pinned by the reference's settled-drape header: sceneHash of the default
scene (src/simulation/scene.ts:438-469) is 4202117243 with 20,930 nodes (public/drape.bin header,
read as data: [0x50415244, 1, 20930, 4202117243])."""
import numpy as np

from bd.material import DENIM
from bd.scene import compile_scene, default_scene_specs, scene_hash, area_code, SEG_BROKEN, SEG_INITIAL_DEFECT
from bd import noise


def test_default_scene_hash_matches_drape_header():
    S = compile_scene(default_scene_specs())
    assert S.node_count == 20930
    assert (S.layers[0].nx, S.layers[0].ny) == (182, 115)
    assert scene_hash(S) == 4202117243
    # initial ruptures are marked as authored defects
    broken = (S.seg_flags & SEG_BROKEN) != 0
    assert np.all((S.seg_flags[broken] & SEG_INITIAL_DEFECT) != 0)


def test_hash32_known_values():
    # lowbias32 of 0 is 0 (every step maps 0 to 0); of 1 a fixed value computed from the definition
    assert int(noise.hash32(0)) == 0
    x = 1
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    assert int(noise.hash32(1)) == x
    # negative inputs wrap like JS ToUint32
    assert int(noise.hash32(-1)) == int(noise.hash32(0xFFFFFFFF))


def test_area_code_rounding():
    assert list(area_code([0.0, 0.5, 1.0, 2.0])) == [1, 128, 255, 255]

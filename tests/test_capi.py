"""CPU checks of the C-ABI library: it loads, exports every symbol include/bd.h declares, and
validates its inputs before touching a device (limits of the reference's GpuSolver constructor,
gpu/solver.ts:298-363, as error codes)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from bd import capi
from bd.configs import sheet_scene
from bd.material import linear_sheet_material

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "bd.h").read_text()
    return sorted(set(re.findall(r"\b(bd_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = capi.load_library()
    names = declared_symbols()
    assert len(names) >= 28
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(capi.EXPORTED) == names


def test_struct_sizes_match_header_layout():
    # bd_yarn: 10 doubles + 2 ints; bd_material: 2 + 2 yarns + 6 doubles
    assert C.sizeof(capi.bd_yarn) == 88
    assert C.sizeof(capi.bd_material) == 2 * 8 + 2 * 88 + 6 * 8
    assert C.sizeof(capi.bd_params) == 8 * 4 + 4 * 2 + 8 + 4 * 2 + 8 + 4 * 10
    assert C.sizeof(capi.bd_stats) == 304 + 8 * capi.BD_MAX_NEWTON + 8
    assert C.sizeof(capi.bd_pick) == 4 + 4 + 3 * 8 + 4 + 4 * 225


def _create(scene, params=None):
    lib = capi.load_library()
    sb = capi.SceneBuffers(scene)
    h = C.c_void_p()
    p = params or capi.make_params()
    rc = lib.bd_create(C.byref(sb.scene), C.byref(p), C.byref(h))
    return rc, lib.bd_last_error(None).decode(), h


def test_create_rejects_lines_beyond_the_limit():
    mat = linear_sheet_material(1e-3, 2e4, 1.6, 1e-5, 0.35, 8e3)
    S = sheet_scene(capi.BD_MAX_LINE + 1, 1e-3, mat)
    rc, msg, _ = _create(S)
    assert rc == -3 and "BD_MAX_LINE" in msg


def test_create_accepts_self_contact_at_a_million_nodes():
    # the reference's reciprocal-bound key holds node + 1 in 20 bits (gpu/selfcontact.ts:84), which
    # refuses a 1024^2 sheet; the 64-bit keys here take it: validation passes and only the missing
    # device stops creation on a CPU host
    mat = linear_sheet_material(1e-3, 2e4, 1.6, 1e-5, 0.35, 8e3)
    S = sheet_scene(1024, 1e-3, mat, perturbation=0.0)
    assert S.node_count == 1 << 20
    rc, msg, h = _create(S, capi.make_params(self_contact=True))
    if rc == 0:
        capi.load_library().bd_destroy(h)
    assert rc != -3, msg
    assert S.node_count <= capi.BD_CONTACT_MAX_NODES


def test_create_rejects_newton_iterations_out_of_range():
    mat = linear_sheet_material(1e-3, 2e4, 1.6, 1e-5, 0.35, 8e3)
    S = sheet_scene(8, 1e-3, mat)
    for n in (0, -1, capi.BD_MAX_NEWTON + 1):
        rc, msg, h = _create(S, capi.make_params(newton_iterations=n))
        if rc == 0:
            capi.load_library().bd_destroy(h)
        assert rc == -1 and "newton_iterations" in msg


def test_create_rejects_inconsistent_scene():
    mat = linear_sheet_material(1e-3, 2e4, 1.6, 1e-5, 0.35, 8e3)
    S = sheet_scene(8, 1e-3, mat)
    S.node_count = 63
    rc, msg, _ = _create(S)
    assert rc == -1 and "node_count" in msg


def test_null_handle_is_an_error():
    lib = capi.load_library()
    assert lib.bd_substep(None, 0, 0) == -1
    assert lib.bd_read(None, 0, 0, None, 0) == -1


def test_graph_and_batch_argument_errors():
    """The capture and batch entry points reject missing handles and empty batches before any device
    work, with the library's message."""
    import ctypes as C
    lib = capi.load_library()
    assert lib.bd_graph_capture(None, 10) == -1
    assert lib.bd_graph_substep(None) == -1
    assert lib.bd_batch_substep(None) == -1
    out = C.c_void_p()
    assert lib.bd_batch_create(None, 0, C.byref(out)) == -1
    assert b"at least one solver" in lib.bd_batch_last_error(None)
    arr = (C.c_void_p * 1)(None)
    assert lib.bd_batch_create(arr, 1, C.byref(out)) == capi.BD_ERR_STATE
    assert b"bd_graph_capture" in lib.bd_batch_last_error(None)
    assert lib.bd_batch_create_concurrent(None, 0, 0, C.byref(out)) == -1
    assert b"at least one solver" in lib.bd_batch_last_error(None)
    assert lib.bd_batch_create_concurrent(arr, 1, -1, C.byref(out)) == -1
    assert lib.bd_batch_create_concurrent(arr, 1, 2, C.byref(out)) == -1
    assert b"null solver" in lib.bd_batch_last_error(None)
    assert lib.bd_batch_destroy(None) == 0

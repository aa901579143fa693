"""The column-major copy used by the v-line passes of long-line levels (lines of 33+ nodes),
against the FP64 oracle with the reference's bars (tests/gpu/correctness.ts): one forward and one
reverse smoother sweep on the fine level (1e-3), the V-cycle (5e-3) and a 2-iteration PCG (2e-2),
on a 136 x 136 hanging sheet whose finer levels keep both layouts and whose coarsest does not."""
import sys

import numpy as np
import pytest

import oracle as O
from bd.capi import BD_BUF_DELTA, BD_BUF_LEVEL_B, BD_BUF_LEVEL_X, BD_BUF_RHS, make_params
from bd.configs import sheet_scene
from bd.material import linear_sheet_material

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_to_scale(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_dual_layout_smoother_vcycle_pcg():
    from bd.solver import GpuSolver
    n = 136
    pitch = 1e-3
    mat = linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / 2.4)
    S = sheet_scene(n, pitch, mat)
    params = make_params(history=False)
    cpu = O.Oracle(S, params)
    for _ in range(4):
        cpu.substep()
    xs = cpu.x.astype(np.float32).astype(np.float64)
    vs = cpu.v.astype(np.float32).astype(np.float64)
    cpu.x[:] = xs
    cpu.v[:] = vs
    say("gpu: create the 136x136 sheet (fine level in both layouts)")
    gpu = GpuSolver(S, params)
    gpu.set_state(xs, vs)
    cpu.predict()
    cpu.assemble()
    gpu.encode_predict()
    gpu.encode_assemble(True)
    assert len(gpu.stats()["level_nodes"]) >= 3
    e = rel_to_scale(gpu.stencil(0).reshape(n * n, -1), cpu.stencil(0).reshape(n * n, -1))
    say(f"fine operator maxAbs/scale = {e:.2e}")
    assert e < 5e-5
    cpu.load_fine_operator(gpu.stencil(0), gpu.fixed(0))
    b = gpu.read(BD_BUF_RHS)
    N = S.node_count
    rng = np.random.default_rng(4)
    x0 = rng.standard_normal(3 * N) * 1e-4
    x0[np.repeat(cpu.fixed(0) == 1, 3)] = 0
    for reverse in (False, True):
        say(f"gpu: one {'reverse' if reverse else 'forward'} smoother sweep from a random iterate")
        gpu.write(BD_BUF_LEVEL_B, b, 0)
        gpu.write(BD_BUF_LEVEL_X, x0, 0)
        gpu.encode_smooth(0, reverse)
        xg = gpu.read(BD_BUF_LEVEL_X, 0)
        xc = cpu.smooth(0, x0.copy(), b, reverse)
        e = rel_to_scale(xg, xc)
        say(f"  sweep maxAbs/scale = {e:.2e}")
        assert e < 1e-3
    say("gpu: V-cycle")
    gpu.write(BD_BUF_LEVEL_B, b, 0)
    gpu.encode_vcycle()
    e = rel_to_scale(gpu.read(BD_BUF_LEVEL_X, 0), cpu.vcycle(b))
    say(f"V-cycle maxAbs/scale = {e:.2e}")
    assert e < 5e-3
    say("gpu: PCG (2 iterations)")
    gpu.write(BD_BUF_RHS, b)
    gpu.encode_pcg()
    e = rel_to_scale(gpu.read(BD_BUF_DELTA), cpu.pcg(b, params.pcg_iterations))
    say(f"PCG step maxAbs/scale = {e:.2e}")
    assert e < 2e-2


def test_autotuned_smoothers_keep_parity():
    """bd_autotune picks a kernel per level and family; the V-cycle and PCG it then runs meet the same
    bars against the oracle (the variants differ only in FP32 summation order)."""
    from bd.solver import GpuSolver
    n = 136
    pitch = 1e-3
    mat = linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / 2.4)
    S = sheet_scene(n, pitch, mat)
    params = make_params(history=False)
    cpu = O.Oracle(S, params)
    for _ in range(4):
        cpu.substep()
    xs = cpu.x.astype(np.float32).astype(np.float64)
    vs = cpu.v.astype(np.float32).astype(np.float64)
    cpu.x[:] = xs
    cpu.v[:] = vs
    gpu = GpuSolver(S, params)
    gpu.set_state(xs, vs)
    cpu.predict()
    gpu.encode_predict()
    gpu.encode_assemble(True)
    chosen = gpu.autotune(2)
    say(f"autotuned smoother kernels (u, v) per level: {chosen}")
    assert all(v in (0, 4, 8) for c in chosen.values() for v in c)
    cpu.load_fine_operator(gpu.stencil(0), gpu.fixed(0))
    b = gpu.read(BD_BUF_RHS)
    gpu.write(BD_BUF_LEVEL_B, b, 0)
    gpu.encode_vcycle()
    e = rel_to_scale(gpu.read(BD_BUF_LEVEL_X, 0), cpu.vcycle(b))
    say(f"V-cycle (autotuned) maxAbs/scale = {e:.2e}")
    assert e < 5e-3
    gpu.write(BD_BUF_RHS, b)
    gpu.encode_pcg()
    e = rel_to_scale(gpu.read(BD_BUF_DELTA), cpu.pcg(b, params.pcg_iterations))
    say(f"PCG (autotuned) maxAbs/scale = {e:.2e}")
    assert e < 2e-2


@pytest.mark.parametrize("sweeps", [2, 3])
def test_repeated_sweeps_keep_parity(sweeps):
    """Repeated pre- and post-smoothing sweeps per level: the V-cycle and PCG meet the same bars
    against the oracle run with the same schedule."""
    from bd.solver import GpuSolver
    n = 136
    pitch = 1e-3
    mat = linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / 2.4)
    S = sheet_scene(n, pitch, mat)
    params = make_params(history=False, smooth_sweeps=sweeps)
    cpu = O.Oracle(S, params)
    for _ in range(4):
        cpu.substep()
    xs = cpu.x.astype(np.float32).astype(np.float64)
    vs = cpu.v.astype(np.float32).astype(np.float64)
    cpu.x[:] = xs
    cpu.v[:] = vs
    say(f"gpu: 136x136 sheet, {sweeps} sweeps")
    gpu = GpuSolver(S, params)
    gpu.set_state(xs, vs)
    cpu.predict()
    cpu.assemble()
    gpu.encode_predict()
    gpu.encode_assemble(True)
    cpu.load_fine_operator(gpu.stencil(0), gpu.fixed(0))
    b = gpu.read(BD_BUF_RHS)
    gpu.write(BD_BUF_LEVEL_B, b, 0)
    gpu.encode_vcycle()
    e = rel_to_scale(gpu.read(BD_BUF_LEVEL_X, 0), cpu.vcycle(b))
    say(f"  V-cycle maxAbs/scale = {e:.2e}")
    assert e < 5e-3
    gpu.write(BD_BUF_RHS, b)
    gpu.encode_pcg()
    e = rel_to_scale(gpu.read(BD_BUF_DELTA), cpu.pcg(b, params.pcg_iterations))
    say(f"  PCG maxAbs/scale = {e:.2e}")
    assert e < 2e-2


def test_every_smoother_variant_keeps_parity():
    """Each smoother kernel variant a level can run (bd_set_smoother), pinned on every level at once:
    one sweep, the V-cycle and PCG meet the same bars against the oracle on a 300 x 300 sheet, whose
    fine lines (300 nodes) take the direct, split and staged kernels."""
    from bd.capi import (BD_SMOOTHER_DIRECT, BD_SMOOTHER_SPLIT, BD_SMOOTHER_SPLIT4, BD_SMOOTHER_FACTORED,
                         BD_SMOOTHER_STAGED4, BD_SMOOTHER_STAGED8, BD_SMOOTHER_DEFAULT)
    from bd.solver import GpuSolver
    n = 300
    pitch = 1e-3
    mat = linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / 2.4)
    S = sheet_scene(n, pitch, mat)
    params = make_params(history=False)
    cpu = O.Oracle(S, params)
    for _ in range(2):
        cpu.substep()
    xs = cpu.x.astype(np.float32).astype(np.float64)
    vs = cpu.v.astype(np.float32).astype(np.float64)
    cpu.x[:] = xs
    cpu.v[:] = vs
    say("gpu: 300x300 sheet")
    gpu = GpuSolver(S, params)
    gpu.set_state(xs, vs)
    cpu.predict()
    cpu.assemble()
    gpu.encode_predict()
    gpu.encode_assemble(True)
    cpu.load_fine_operator(gpu.stencil(0), gpu.fixed(0))
    b = gpu.read(BD_BUF_RHS)
    rng = np.random.default_rng(6)
    x0 = rng.standard_normal(3 * S.node_count) * 1e-4
    x0[np.repeat(cpu.fixed(0) == 1, 3)] = 0
    xs_ref = cpu.smooth(0, x0.copy(), b, False)
    xv_ref = cpu.vcycle(b)
    xp_ref = cpu.pcg(b, params.pcg_iterations)
    nlev = len(gpu.stats()["level_nodes"])
    for variant in (BD_SMOOTHER_DIRECT, BD_SMOOTHER_SPLIT, BD_SMOOTHER_SPLIT4, BD_SMOOTHER_FACTORED, BD_SMOOTHER_STAGED4,
                    BD_SMOOTHER_STAGED8):
        pinned = []
        for k in range(nlev):
            for fam in (0, 1):
                try:
                    gpu.set_smoother(k, fam, variant)
                    pinned.append((k, fam))
                except RuntimeError:
                    gpu.set_smoother(k, fam, BD_SMOOTHER_DEFAULT)
        if variant in (BD_SMOOTHER_SPLIT, BD_SMOOTHER_SPLIT4, BD_SMOOTHER_FACTORED):
            assert (0, 0) in pinned and (0, 1) in pinned, "the fine level keeps a rows buffer"
        say(f"variant {variant}: pinned on {pinned}")
        gpu.write(BD_BUF_LEVEL_B, b, 0)
        gpu.write(BD_BUF_LEVEL_X, x0, 0)
        gpu.encode_smooth(0, False)
        e = rel_to_scale(gpu.read(BD_BUF_LEVEL_X, 0), xs_ref)
        say(f"  sweep maxAbs/scale = {e:.2e}")
        assert e < 1e-3
        gpu.write(BD_BUF_LEVEL_B, b, 0)
        gpu.encode_vcycle()
        e = rel_to_scale(gpu.read(BD_BUF_LEVEL_X, 0), xv_ref)
        say(f"  V-cycle maxAbs/scale = {e:.2e}")
        assert e < 5e-3
        gpu.write(BD_BUF_RHS, b)
        gpu.encode_pcg()
        e = rel_to_scale(gpu.read(BD_BUF_DELTA), xp_ref)
        say(f"  PCG maxAbs/scale = {e:.2e}")
        assert e < 2e-2


@pytest.mark.parametrize("n", [136, 300])
def test_fp16_vcycle_keeps_parity(n):
    """BD_PRECISION_FP16: the V-cycle reads every coarse level's operator as FP16 at the level's
    power-of-two scale (FP16 pairs; the fine level stays FP32). A fine and a level-1 sweep, the V-cycle
    and a 2-iteration PCG meet the FP32 bars against the oracle run with the same rounding
    (oracle.fp16_round, pinned to IEEE binary16)."""
    from bd.capi import BD_PRECISION_FP16
    from bd.solver import GpuSolver
    pitch = 1e-3
    mat = linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / 2.4)
    S = sheet_scene(n, pitch, mat)
    params = make_params(history=False, precision=BD_PRECISION_FP16)
    cpu = O.Oracle(S, params)
    for _ in range(2):
        cpu.substep()
    xs = cpu.x.astype(np.float32).astype(np.float64)
    vs = cpu.v.astype(np.float32).astype(np.float64)
    cpu.x[:] = xs
    cpu.v[:] = vs
    say(f"gpu: {n}x{n} sheet, FP16 V-cycle operators")
    gpu = GpuSolver(S, params)
    gpu.set_state(xs, vs)
    cpu.predict()
    cpu.assemble()
    gpu.encode_predict()
    gpu.encode_assemble(True)
    cpu.load_fine_operator(gpu.stencil(0), gpu.fixed(0))
    b = gpu.read(BD_BUF_RHS)
    rng = np.random.default_rng(8)
    x0 = rng.standard_normal(3 * S.node_count) * 1e-4
    x0[np.repeat(cpu.fixed(0) == 1, 3)] = 0
    gpu.write(BD_BUF_LEVEL_B, b, 0)
    gpu.write(BD_BUF_LEVEL_X, x0, 0)
    gpu.encode_smooth(0, False)
    e = rel_to_scale(gpu.read(BD_BUF_LEVEL_X, 0), cpu.smooth(0, x0.copy(), b, False))
    say(f"  sweep maxAbs/scale = {e:.2e}")
    assert e < 1e-3
    n1 = gpu.stats()["level_nodes"][1]
    b1 = rng.standard_normal(3 * n1)
    x1 = rng.standard_normal(3 * n1) * 1e-2
    gpu.write(BD_BUF_LEVEL_B, b1, 1)
    gpu.write(BD_BUF_LEVEL_X, x1, 1)
    gpu.encode_smooth(1, False)
    xg1 = gpu.read(BD_BUF_LEVEL_X, 1)
    xc1 = cpu.smooth(1, np.where(np.repeat(cpu.fixed(1) == 1, 3), 0.0, x1), np.where(np.repeat(cpu.fixed(1) == 1, 3), 0.0, b1), False)
    free1 = ~np.repeat(cpu.fixed(1) == 1, 3)
    e = rel_to_scale(xg1[free1], xc1[free1])
    say(f"  level-1 FP16 sweep maxAbs/scale = {e:.2e}")
    assert e < 1e-3
    gpu.write(BD_BUF_LEVEL_B, b, 0)
    gpu.encode_vcycle()
    e = rel_to_scale(gpu.read(BD_BUF_LEVEL_X, 0), cpu.vcycle(b))
    say(f"  V-cycle maxAbs/scale = {e:.2e}")
    assert e < 5e-3
    gpu.write(BD_BUF_RHS, b)
    gpu.encode_pcg()
    e = rel_to_scale(gpu.read(BD_BUF_DELTA), cpu.pcg(b, params.pcg_iterations))
    say(f"  PCG maxAbs/scale = {e:.2e}")
    assert e < 2e-2


def test_fp16_pcg_reaches_the_fp32_tolerance():
    """The FP16 coarse levels only precondition: PCG in tolerance mode (FP32 matvec, FP64 residual
    replacement) reaches the same relative residual as with FP32 operators, in about as many
    iterations (on the 1024^2 workload: 124.2 against 123.4 per frame)."""
    from bd.capi import BD_PRECISION_FP16
    from bd.solver import GpuSolver
    n = 200
    pitch = 1e-3
    mat = linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / 2.4)
    S = sheet_scene(n, pitch, mat)
    its = {}
    for prec in (0, BD_PRECISION_FP16):
        gpu = GpuSolver(S, make_params(history=False, precision=prec))
        gpu.encode_predict()
        gpu.encode_assemble(True)
        gpu.set_pcg(2, 1e-6, 400)
        gpu.encode_pcg()
        st = gpu.stats()
        its[prec] = st["pcg_iterations"][0]
        say(f"precision {prec}: {its[prec]} iterations, reported rel-res {st['rel_residual'][0]:.2e}")
        assert st["rel_residual"][0] <= 1e-6
        gpu.close()
    assert its[BD_PRECISION_FP16] <= 1.1 * its[0] + 2


def test_set_smoother_rejects_what_a_level_cannot_run():
    """bd_set_smoother: unknown levels, families and variants, and the split variant on a level without
    a rows buffer, are rejected with BD_ERR_INVALID; a valid choice pins and the default unpins."""
    from bd.capi import BD_SMOOTHER_DEFAULT, BD_SMOOTHER_SPLIT, BD_SMOOTHER_STAGED8
    from bd.solver import GpuSolver
    pitch = 1e-3
    mat = linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / 2.4)
    S = sheet_scene(136, pitch, mat)  # lines of 136 nodes: no level keeps a rows buffer
    gpu = GpuSolver(S, make_params(history=False))
    with pytest.raises(RuntimeError):
        gpu.set_smoother(0, 0, BD_SMOOTHER_STAGED8)  # before the first assembly
    gpu.encode_predict()
    gpu.encode_assemble(True)
    nlev = len(gpu.stats()["level_nodes"])
    for level, fam, variant in ((nlev, 0, 0), (-1, 0, 0), (0, 2, 0), (0, 0, 5), (0, 0, BD_SMOOTHER_SPLIT)):
        with pytest.raises(RuntimeError):
            gpu.set_smoother(level, fam, variant)
    gpu.set_smoother(0, 1, BD_SMOOTHER_STAGED8)
    gpu.set_smoother(0, 1, BD_SMOOTHER_DEFAULT)
    gpu.encode_vcycle()
    assert np.isfinite(gpu.read(BD_BUF_LEVEL_X, 0)).all()
    gpu.close()


def test_one_colour_coarse_levels_keep_parity():
    """bd_set_line_colors(level, 1) on every Galerkin level of a 136 x 136 sheet (each family's lines
    solved at once against the lagged iterate, the iterate ping-ponged between two buffers), against
    the oracle smoothing the same levels with one colour: a level-1 sweep in both directions, the
    V-cycle and a 2-iteration PCG at the usual bars; then a tolerance-mode PCG reaches rel-res 1e-6,
    confirmed in FP64, and 3 colours can be set back."""
    from bd.solver import GpuSolver
    n = 136
    pitch = 1e-3
    mat = linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / 2.4)
    S = sheet_scene(n, pitch, mat)
    params = make_params(history=False)
    cpu = O.Oracle(S, params)
    for _ in range(4):
        cpu.substep()
    xs = cpu.x.astype(np.float32).astype(np.float64)
    vs = cpu.v.astype(np.float32).astype(np.float64)
    cpu.x[:] = xs
    cpu.v[:] = vs
    say("gpu: 136x136 sheet, one line colour on the Galerkin levels")
    gpu = GpuSolver(S, params)
    gpu.set_state(xs, vs)
    cpu.predict()
    cpu.assemble()
    gpu.encode_predict()
    gpu.encode_assemble(True)
    nlev = len(gpu.stats()["level_nodes"])
    assert nlev >= 3
    for k in range(1, nlev):
        gpu.set_line_colors(k, 1)
        cpu.set_line_colors(k, 1)
    with pytest.raises(RuntimeError, match="1, 2 or 3"):
        gpu.set_line_colors(1, 4)
    with pytest.raises(RuntimeError, match="split kernels"):
        gpu.set_line_colors(1, 2)
    cpu.load_fine_operator(gpu.stencil(0), gpu.fixed(0))
    b = gpu.read(BD_BUF_RHS)
    # a level-1 sweep on the Galerkin operator both sides hold
    cn = cpu.level_nodes(1)
    rng = np.random.default_rng(8)
    b1 = rng.standard_normal(3 * cn) * 1e-3
    x1 = rng.standard_normal(3 * cn) * 1e-6
    fx1 = np.repeat(cpu.fixed(1) == 1, 3)
    b1[fx1] = 0
    x1[fx1] = 0
    for reverse in (False, True):
        gpu.write(BD_BUF_LEVEL_B, b1, 1)
        gpu.write(BD_BUF_LEVEL_X, x1, 1)
        gpu.encode_smooth(1, reverse)
        e = rel_to_scale(gpu.read(BD_BUF_LEVEL_X, 1), cpu.smooth(1, x1.copy(), b1, reverse))
        say(f"  level-1 one-colour sweep ({'reverse' if reverse else 'forward'}) maxAbs/scale = {e:.2e}")
        assert e < 1e-3
    gpu.write(BD_BUF_LEVEL_B, b, 0)
    gpu.encode_vcycle()
    e = rel_to_scale(gpu.read(BD_BUF_LEVEL_X, 0), cpu.vcycle(b))
    say(f"V-cycle maxAbs/scale = {e:.2e}")
    assert e < 5e-3
    gpu.write(BD_BUF_RHS, b)
    gpu.encode_pcg()
    e = rel_to_scale(gpu.read(BD_BUF_DELTA), cpu.pcg(b, params.pcg_iterations))
    say(f"PCG step maxAbs/scale = {e:.2e}")
    assert e < 2e-2
    gpu.set_pcg(0, 1e-6, 300)
    gpu.write(BD_BUF_RHS, b)
    gpu.encode_pcg()
    st = gpu.stats()
    say(f"tolerance PCG: {st['pcg_iterations'][0]} iterations, rel-res {st['rel_residual'][0]:.2e}")
    assert st["rel_residual"][0] <= 1.05e-6
    for k in range(1, nlev):
        gpu.set_line_colors(k, 3)
    gpu.write(BD_BUF_LEVEL_B, b, 0)
    gpu.encode_vcycle()
    for k in range(1, nlev):
        cpu.set_line_colors(k, 3)
    e = rel_to_scale(gpu.read(BD_BUF_LEVEL_X, 0), cpu.vcycle(b))
    say(f"V-cycle with 3 colours again maxAbs/scale = {e:.2e}")
    assert e < 5e-3
    gpu.close()


def test_batch_copy_level_keeps_parity():
    """A 520 x 520 sheet's fine level (lines of 520 nodes) keeps x in both layouts but refreshes the
    other layout once per family's passes (one tiled copy) instead of in every pass. With the direct,
    four-row split and factored kernels pinned, and with 3, 2 and 1 line colours (two colours: lines
    two apart together, their couplings majorised, split kernels only; the one-colour sweep ping-pongs
    between the pair of buffers), a forward and a reverse sweep, the V-cycle and a 2-iteration PCG meet
    the usual bars against the oracle."""
    from bd.capi import BD_SMOOTHER_DIRECT, BD_SMOOTHER_FACTORED, BD_SMOOTHER_SPLIT4
    from bd.solver import GpuSolver
    n = 520
    pitch = 1e-3
    mat = linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / 2.4)
    S = sheet_scene(n, pitch, mat)
    params = make_params(history=False)
    cpu = O.Oracle(S, params)
    cpu.substep()
    xs = cpu.x.astype(np.float32).astype(np.float64)
    vs = cpu.v.astype(np.float32).astype(np.float64)
    cpu.x[:] = xs
    cpu.v[:] = vs
    say("gpu: 520x520 sheet")
    gpu = GpuSolver(S, params)
    gpu.set_state(xs, vs)
    cpu.predict()
    cpu.assemble()
    gpu.encode_predict()
    gpu.encode_assemble(True)
    cpu.load_fine_operator(gpu.stencil(0), gpu.fixed(0))
    b = gpu.read(BD_BUF_RHS)
    rng = np.random.default_rng(9)
    x0 = rng.standard_normal(3 * S.node_count) * 1e-4
    x0[np.repeat(cpu.fixed(0) == 1, 3)] = 0
    for colours in (3, 2, 1):
        gpu.set_line_colors(0, colours)
        cpu.set_line_colors(0, colours)
        refs = {rev: cpu.smooth(0, x0.copy(), b, rev) for rev in (False, True)}
        xv_ref = cpu.vcycle(b)
        variants = (BD_SMOOTHER_SPLIT4, BD_SMOOTHER_FACTORED) if colours == 2 else (
            BD_SMOOTHER_DIRECT, BD_SMOOTHER_SPLIT4, BD_SMOOTHER_FACTORED)
        if colours == 2:
            with pytest.raises(RuntimeError, match="cannot run"):
                gpu.set_smoother(0, 0, BD_SMOOTHER_DIRECT)
        for variant in variants:
            for fam in (0, 1):
                gpu.set_smoother(0, fam, variant)
            say(f"{colours} colour(s), variant {variant}")
            for rev in (False, True):
                gpu.write(BD_BUF_LEVEL_B, b, 0)
                gpu.write(BD_BUF_LEVEL_X, x0, 0)
                gpu.encode_smooth(0, rev)
                e = rel_to_scale(gpu.read(BD_BUF_LEVEL_X, 0), refs[rev])
                say(f"  sweep ({'reverse' if rev else 'forward'}) maxAbs/scale = {e:.2e}")
                assert e < 1e-3
            gpu.write(BD_BUF_LEVEL_B, b, 0)
            gpu.encode_vcycle()
            e = rel_to_scale(gpu.read(BD_BUF_LEVEL_X, 0), xv_ref)
            say(f"  V-cycle maxAbs/scale = {e:.2e}")
            assert e < 5e-3
        gpu.write(BD_BUF_RHS, b)
        gpu.encode_pcg()
        e = rel_to_scale(gpu.read(BD_BUF_DELTA), cpu.pcg(b, params.pcg_iterations))
        say(f"  PCG maxAbs/scale = {e:.2e}")
        assert e < 2e-2
    gpu.close()

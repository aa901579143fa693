"""Oracle multigrid and solver pinned by the reference's unit tests
(tests/unit/newtonMultigrid.test.ts, referenceSolver.test.ts) and brute force on tiny inputs."""
import math

import numpy as np
import pytest

import oracle as O
from bd.capi import make_params
from bd.material import DENIM
from bd.scene import compile_scene, coupon, SEG_BROKEN
from conftest import lcg


def ref_params(h=1 / 240, gravity=(0, 0, -9.81), newton=2, pcg=4):
    """SolverParams with MG_DEFAULTS (newtonMultigrid.ts:36): no lagging, no line search."""
    return make_params(h=h, gravity=gravity, newton_iterations=newton, pcg_iterations=pcg, coarse_sweeps=6,
                       max_coarsest_nodes=48, lag_coarse=False, line_search=False, history=True)


def mg_coupon(nu, nv):  # newtonMultigrid.test.ts:27-43
    return coupon(nu, nv, angle=0.3, compressU=0.02, compressV=0.01, shear=0.01, collarRows=2, seed=3,
                  perturbation=0.3e-3)


def rs_coupon(nu, nv, **kw):  # referenceSolver.test.ts:7-24
    spec = dict(angle=0.3, compressU=0.01, compressV=0.005, shear=0.01, collarRows=1, seed=7, perturbation=0.2e-3)
    spec.update(kw)
    return coupon(nu, nv, **spec)


def to_dense(A, nx, ny):
    N = A.shape[0]
    M = np.zeros((3 * N, 3 * N))
    for n in range(N):
        i, j = n % nx, n // nx
        for dj in range(-2, 3):
            for di in range(-2, 3):
                if 0 <= i + di < nx and 0 <= j + dj < ny:
                    m = n + dj * nx + di
                    M[3 * n:3 * n + 3, 3 * m:3 * m + 3] = A[n, (dj + 2) * 5 + di + 2]
    return M


def test_solve_penta_residual():  # newtonMultigrid.test.ts:62-93
    r = lcg(9)
    n = 9
    D = np.zeros((n, 3, 3))
    E = np.zeros((n, 3, 3))
    F = np.zeros((n, 3, 3))
    for k in range(n):
        for q in range(9):
            E[k].flat[q] = r()
            F[k].flat[q] = 0.5 * r()
        for a in range(3):
            for b in range(3):
                D[k, a, b] = 12 if a == b else 0.3 * (a + b - 2)
    rhs = np.array([r() for _ in range(3 * n)])
    x = O.solve_penta(D, E, F, rhs).reshape(n, 3)
    for k in range(n):
        s = D[k] @ x[k]
        if k + 1 < n: s += E[k] @ x[k + 1]
        if k + 2 < n: s += F[k] @ x[k + 2]
        if k >= 1: s += E[k - 1].T @ x[k - 1]
        if k >= 2: s += F[k - 2].T @ x[k - 2]
        assert np.max(np.abs(s - rhs[3 * k:3 * k + 3])) < 1e-12


def test_operator_symmetric_and_galerkin_equals_ptap():  # newtonMultigrid.test.ts:95-151
    S = compile_scene([mg_coupon(11, 9)])
    L = S.layers[0]
    o = O.Oracle(S, ref_params())
    o.predict()
    o.assemble()
    o.build_hierarchy()
    A = o.stencil(0)
    M = to_dense(A, L.nx, L.ny)
    assert np.max(np.abs(M - M.T)) / np.max(np.abs(M)) < 1e-12
    assert o.level_count() >= 2
    fixed = o.fixed(0)
    idx, w = o.prolongation(0)
    nc = o.level_nodes(1)
    Pd = np.zeros((3 * A.shape[0], 3 * nc))
    for f in range(A.shape[0]):
        for k in range(4):
            if idx[f, k] >= 0:
                for q in range(3):
                    Pd[3 * f + q, 3 * idx[f, k] + q] = w[f, k]
    Mf = M.copy()
    for n in np.nonzero(fixed)[0]:
        Mf[3 * n:3 * n + 3, :] = 0
        Mf[:, 3 * n:3 * n + 3] = 0
    G = Pd.T @ Mf @ Pd
    cx, cy = (L.nx + 1) // 2 + (0 if L.nx % 2 else 1), None
    Cst = o.stencil(1)
    cfixed = o.fixed(1)
    cnx = math.ceil((L.nx - 1) / 2) + 1
    cny = math.ceil((L.ny - 1) / 2) + 1
    Mc = to_dense(Cst, cnx, cny)
    free = np.repeat(cfixed == 0, 3)
    err = np.max(np.abs((G - Mc)[np.ix_(free, free)]))
    assert err / np.max(np.abs(G[np.ix_(free, free)])) < 1e-10


def test_smoothing_vcycle_pcg_reduce_residual():  # newtonMultigrid.test.ts:153-194
    S = compile_scene([mg_coupon(40, 30)])
    o = O.Oracle(S, ref_params())
    o.predict()
    o.assemble()
    o.build_hierarchy()
    fixed = o.fixed(0)
    N = S.node_count
    r = lcg(4)
    b = np.array([r() for _ in range(3 * N)])
    b.reshape(N, 3)[fixed == 1] = 0
    norm = lambda x: np.linalg.norm(b - o.apply(0, x))
    r0 = norm(np.zeros(3 * N))
    x = o.smooth(0, np.zeros(3 * N), b)
    assert norm(x) < r0
    assert o.level_count() > 2
    y = o.vcycle(b)
    assert norm(y) < 0.5 * r0
    xs = np.array([1e-5 * r() for _ in range(3 * N)])
    xs.reshape(N, 3)[fixed == 1] = 0
    bs = o.apply(0, xs)

    def enorm(x):
        e = x - xs
        return math.sqrt(e @ o.apply(0, e))

    e0 = enorm(np.zeros(3 * N))
    z = o.pcg(bs, 3)
    assert enorm(z) / e0 < 5e-3


def dense_smooth(A, fixed, nx, ny, x0, b, reverse=False, colors=3):
    """smoothLines (newtonMultigrid.ts:659-733) restated with a dense solve of every line system."""
    N = nx * ny
    x = x0.reshape(N, 3).copy()
    order = [(f, c) for f in (1, 0) for c in range(colors)]
    if reverse:
        order.reverse()
    for family, color in order:
        xo = x.copy()
        lines = ny if family == 0 else nx
        for line in range(color, lines, colors):
            nodes = [line * nx + k for k in range(nx)] if family == 0 else [k * nx + line for k in range(ny)]
            n3 = 3 * len(nodes)
            M = np.zeros((n3, n3))
            rhs = np.zeros(n3)
            for k, n in enumerate(nodes):
                rhs[3 * k:3 * k + 3] = b[3 * n:3 * n + 3]
                if fixed[n]:
                    M[3 * k:3 * k + 3, 3 * k:3 * k + 3] = np.eye(3)
                    continue
                maj = 0.0
                i, j = n % nx, n // nx
                for dj in range(-2, 3):
                    for di in range(-2, 3):
                        if not (0 <= i + di < nx and 0 <= j + dj < ny):
                            continue
                        m = n + dj * nx + di
                        if fixed[m] and not (di == 0 and dj == 0):
                            continue
                        B = A[n, (dj + 2) * 5 + di + 2]
                        along, across = (di, dj) if family == 0 else (dj, di)
                        if across == 0 and abs(along) <= 1:
                            if along == 1 or along == 0:
                                M[3 * k:3 * k + 3, 3 * (k + along):3 * (k + along) + 3] += B
                            if along == 1:
                                M[3 * (k + 1):3 * (k + 1) + 3, 3 * k:3 * k + 3] += B.T
                            continue
                        # newtonMultigrid.ts:705 (2 and 3 colours); 1 colour majorises along the line only
                        if across == 0 or (colors > 1 and abs(across) == colors):
                            maj += np.linalg.norm(B)
                        rhs[3 * k:3 * k + 3] -= B @ xo[m]
                M[3 * k:3 * k + 3, 3 * k:3 * k + 3] += maj * np.eye(3)
                rhs[3 * k:3 * k + 3] += maj * xo[n]
            sol = np.linalg.solve(M, rhs).reshape(-1, 3)
            for k, n in enumerate(nodes):
                if not fixed[n]:
                    x[n] = sol[k]
    return x.ravel()


@pytest.mark.parametrize("colors", [3, 2, 1])
@pytest.mark.parametrize("reverse", [False, True])
def test_smoother_matches_dense_line_solves(reverse, colors):
    """Brute force on a small coupon: a full colour-ordered sweep equals dense solves of every line's
    block-tridiagonal system (lower blocks = transposes, +-2 along-line couplings majorised by their
    Frobenius norms; with 2 colours the couplings of lines two apart too), each line against the
    iterate its colour started from, up to the 2e-6 relative pivot shift; for the reference's 3 and 2
    colours per family and for 1 (every line of a family at once)."""
    S = compile_scene([mg_coupon(10, 8)])
    L = S.layers[0]
    o = O.Oracle(S, ref_params())
    o.set_line_colors(0, colors)
    o.predict()
    o.assemble()
    A = o.stencil(0)
    fixed = o.fixed(0)
    N = S.node_count
    r = lcg(11)
    b = np.array([r() for _ in range(3 * N)])
    x0 = np.array([1e-3 * r() for _ in range(3 * N)])
    x0.reshape(N, 3)[fixed == 1] = 0
    x = o.smooth(0, x0, b, reverse)
    ref = dense_smooth(A, fixed, L.nx, L.ny, x0, b, reverse, colors)
    assert np.max(np.abs(x - ref)) <= 1e-4 * np.max(np.abs(ref))
    # a plausible slip (dropping the majorisation) is caught at this tolerance
    assert np.max(np.abs(x - x0)) > 1e-3 * np.max(np.abs(ref))


def test_fd_gradient():  # referenceSolver.test.ts:39-64
    S = compile_scene([rs_coupon(9, 8)])
    o = O.Oracle(S, ref_params())
    v = o.v
    v[:] = 0.01 * np.sin(np.arange(len(v)) * 1.7)
    o.predict()
    rnd = lcg(1)
    dx = o.dx.reshape(-1, 3)
    for n in range(S.node_count):
        if S.pinned[n]:
            continue
        for q in range(3):
            dx[n, q] = 2e-5 * rnd()
    free = [n for n in range(S.node_count) if not S.pinned[n]][:20]
    dxf = o.dx
    for n in free:
        g = o.node_gradient(n)
        for q in range(3):
            h = 1e-9
            k = 3 * n + q
            x0 = dxf[k]
            dxf[k] = x0 + h
            ep = o.potential()
            dxf[k] = x0 - h
            em = o.potential()
            dxf[k] = x0
            fdv = (ep - em) / (2 * h)
            scale = max(1e-6, abs(fdv), abs(g[q]))
            assert abs(fdv - g[q]) / scale < 2e-4


def test_rigid_motion_invariance():  # referenceSolver.test.ts:66-100
    S = compile_scene([rs_coupon(8, 7, compressU=0.0, compressV=0.0, shear=0)])
    o = O.Oracle(S, ref_params())
    rnd = lcg(5)
    dx = o.dx.reshape(-1, 3)
    for n in range(S.node_count):
        if S.pinned[n]:
            continue
        for q in range(3):
            dx[n, q] = 3e-5 * rnd()
    o.pred[:] = o.dx
    e0 = o.potential()
    ax = np.array([0.3, -0.5, 0.81])
    kx, ky, kz = ax / np.linalg.norm(ax)
    th = 0.7
    c, sn = math.cos(th), math.sin(th)
    R = np.array([[c + kx * kx * (1 - c), kx * ky * (1 - c) - kz * sn, kx * kz * (1 - c) + ky * sn],
                  [ky * kx * (1 - c) + kz * sn, c + ky * ky * (1 - c), ky * kz * (1 - c) - kx * sn],
                  [kz * kx * (1 - c) - ky * sn, kz * ky * (1 - c) + kx * sn, c + kz * kz * (1 - c)]])
    t = np.array([0.01, -0.02, 0.005])
    x = o.x.reshape(-1, 3)
    y = x + dx
    rx = x @ R.T + t
    ry = y @ R.T + t
    x[:] = rx
    dx[:] = ry - rx
    o.pred[:] = o.dx
    e1 = o.potential()
    assert abs(e1 - e0) / max(e0, 1e-12) < 1e-7


def test_newton_decreases_potential():  # referenceSolver.test.ts:103-118
    S = compile_scene([rs_coupon(30, 24)])
    o = O.Oracle(S, ref_params())
    for _ in range(10):
        o.substep()
    o.predict()
    e0 = o.potential()
    prev = e0
    for it in range(4):
        o.newton_step(True)
        e = o.potential()
        assert e <= prev * 1.02
        prev = e
    assert prev < e0


def test_settles_under_gravity():  # referenceSolver.test.ts:120-135
    S = compile_scene([rs_coupon(20, 16)])
    o = O.Oracle(S, ref_params())
    for _ in range(240):
        o.substep()
    v = o.v.reshape(-1, 3)
    x = o.x.reshape(-1, 3)
    assert np.all(np.isfinite(x))
    assert np.max(np.linalg.norm(v, axis=1)) < 0.02
    assert np.max(np.abs(x[:, 2])) < 0.005
    assert o.broken_count == 0


def test_pull_ruptures_and_conserves_yarn_length():  # referenceSolver.test.ts:137-160
    S = compile_scene([rs_coupon(14, 10, compressU=0, compressV=0, shear=0, perturbation=0)])
    o = O.Oracle(S, ref_params(gravity=(0, 0, 0)))
    L0 = o.yarn_material_length()
    L = S.layers[0]
    top = [L.node_offset + (L.ny - 1) * L.nx + i for i in range(L.nx)]
    vdir = np.array([L.mapA[1], L.mapA[3]])
    vdir /= np.linalg.norm(vdir)
    x = o.x.reshape(-1, 3)
    k = 0
    while k < 400 and o.broken_count < 6:
        for n in top:
            x[n, 0] += (0.02 / 240) * vdir[0]
            x[n, 1] += (0.02 / 240) * vdir[1]
        o.substep()
        k += 1
    assert o.broken_count > 0
    assert int(np.count_nonzero(o.seg_flags & SEG_BROKEN)) >= o.broken_count
    assert abs(o.yarn_material_length() - L0) / L0 < 1e-9


def test_exact_coarse_solve_is_the_inverse():
    """BD_COARSE_EXACT (an extension): with a single level the V-cycle is the banded Cholesky solve,
    which must agree with a dense solve of the same operator; with a hierarchy it is the exact
    two-grid correction at the coarsest level, so the V-cycle beats the swept one."""
    S = compile_scene([mg_coupon(11, 9)])
    p = ref_params()
    p.coarse_solve = 1
    p.max_coarsest_nodes = 10 ** 6   # the fine level is the coarsest
    o = O.Oracle(S, p)
    o.predict()
    o.assemble()
    o.build_hierarchy()
    assert o.level_count() == 1
    N = S.node_count
    A = to_dense(o.stencil(0), 11, 9)
    r = lcg(9)
    b = np.array([r() for _ in range(3 * N)])
    b.reshape(N, 3)[o.fixed(0) == 1] = 0
    x = o.vcycle(b)
    xd = np.linalg.solve(A, b)
    assert np.max(np.abs(x - xd)) <= 1e-10 * np.max(np.abs(xd))
    # on a hierarchy: the exact coarsest solve is the limit of many coarsest sweeps, up to the
    # sweeps' shifted pivots (PIVOT_SHIFT 2e-6 of each diagonal moves their fixed point by ~1e-4)
    S2 = compile_scene([mg_coupon(40, 30)])
    out = []
    for mode, sweeps in ((0, 400), (1, 2)):
        q = ref_params()
        q.coarse_solve = mode
        q.coarse_sweeps = sweeps
        o2 = O.Oracle(S2, q)
        o2.predict()
        o2.assemble()
        o2.build_hierarchy()
        assert o2.level_count() > 2
        N2 = S2.node_count
        r2 = lcg(4)
        b2 = np.array([r2() for _ in range(3 * N2)])
        b2.reshape(N2, 3)[o2.fixed(0) == 1] = 0
        out.append(o2.vcycle(b2))
    assert np.max(np.abs(out[1] - out[0])) <= 5e-4 * np.max(np.abs(out[0]))


def test_lagged_exact_coarse_solve():
    """BD_COARSE_EXACT_LAGGED (an extension): a hierarchy build that refactors (outside a substep, or a
    substep's first Newton iteration) gives the BD_COARSE_EXACT V-cycle bit for bit; a later Newton
    iteration's build keeps the first factorization, so its V-cycle differs from the refactored one
    by the change of the coarsest operator, and a one-iteration substep is BD_COARSE_EXACT's."""
    S = compile_scene([mg_coupon(40, 30)])
    N = S.node_count
    r = lcg(5)
    b = np.array([r() for _ in range(3 * N)])
    runs = {}
    for mode in (1, 2):
        q = ref_params()
        q.coarse_solve = mode
        o = O.Oracle(S, q)
        o.predict()
        o.newton_step(True)       # refactors in both modes
        bb = b.copy()
        bb.reshape(N, 3)[o.fixed(0) == 1] = 0
        first = o.vcycle(bb)
        o.newton_step(False)      # the lagged mode keeps the factorization of the step before
        later = o.vcycle(bb)
        o.build_hierarchy()       # outside a substep: refactors in both modes
        again = o.vcycle(bb)
        runs[mode] = (first, later, again, o.x.copy())
    (e1, e2, e3, _), (l1, l2, l3, _) = runs[1], runs[2]
    assert np.array_equal(e1, l1)
    assert np.array_equal(e3, l3)
    d = np.max(np.abs(l2 - e2)) / np.max(np.abs(e2))
    assert 0 < d < 1e-2
    one = {}
    for mode in (1, 2):
        q = ref_params()
        q.coarse_solve = mode
        q.newton_iterations = 1
        q.newton_max = 1
        o = O.Oracle(S, q)
        for _ in range(3):
            o.substep()
        one[mode] = o.x.copy()
    assert np.array_equal(one[1], one[2])


def test_fp16_rounding_matches_ieee_half():
    """The oracle's FP16 rounding of the V-cycle operators (BD_PRECISION_FP16) is IEEE binary16 round
    to nearest even (numpy's float32 -> float16 conversion), subnormals included, saturated at the
    largest finite half instead of overflowing to inf."""
    rng = np.random.default_rng(11)
    vals = np.concatenate([
        rng.standard_normal(4000) * 10.0 ** rng.uniform(-9, 4.8, 4000),
        np.array([0.0, -0.0, 6.103515625e-05, 6.0e-05, 5.96e-08, 2.98e-08, 2.99e-08, 1.0, 1.0 + 2 ** -11,
                  1.0 + 3 * 2 ** -11, 2048.0 + 1.0, 2048.0 + 3.0, 65504.0, 65519.0, 70000.0, -1e6]),
    ])
    got = np.array([O.fp16_round(v) for v in vals])
    want = np.clip(vals.astype(np.float32), -65504, 65504).astype(np.float16).astype(np.float64)
    assert np.array_equal(got, want), vals[got != want][:5]


def _linear_sheet(n, pitch=1e-3, bend_k=1e-5, shear=2e4 / 2.4, drag=None):
    from bd.configs import sheet_scene
    from bd.material import linear_sheet_material
    mat = linear_sheet_material(pitch, 2e4, 1.6, bend_k, 0.35, shear)
    if drag is not None:
        mat["airDrag"] = drag
    return sheet_scene(n, pitch, mat, perturbation=0.3e-3)


def test_energy_change_equals_potential_differences():
    # energyChange (solver.ts:430-534) against potential (solver.ts:387-422 integrand) on a linear-law
    # sheet: the axial and shear trapezoids are exact for linear laws, bendPsiDelta is exact by
    # construction and every node term is quadratic, so each of the four alphas must reproduce
    # Phi(dx + alpha d) - Phi(dx) to rounding. A dropped term, a wrong alpha power or a sign error
    # shows up at O(1) relative.
    S = _linear_sheet(12)
    grip = {"nodes": [0, 1, 12, 13], "weights": [0.25] * 4, "offsets": np.zeros((4, 3)),
            "goal": S.position[[0, 1, 12, 13]].mean(0) + np.array([2e-3, -1e-3, 0]), "k": 2e4}
    o = O.Oracle(S, make_params(h=1 / 120, newton_iterations=1, line_search=True, history=False), [grip])
    o.predict()
    r = np.random.default_rng(5)
    free = np.repeat(o.pinned == 0, 3)
    # steps of a few percent of the 1 mm pitch keep the shear diagonals out of their lock ramp
    # (shearLaw is linear while compression stays under the jam strain, laws.ts:96-115)
    o.dx[:] = np.where(free, 2e-5 * r.standard_normal(3 * S.node_count), 0.0)
    d = np.where(free, 4e-5 * r.standard_normal(3 * S.node_count), 0.0)
    dx0 = o.dx.copy()
    e0 = o.potential()
    e, mag = o.energy_change(d)
    for k, a in enumerate((1.0, 0.5, 0.25, 0.125)):
        o.dx[:] = dx0 + a * d
        diff = o.potential() - e0
        o.dx[:] = dx0
        assert abs(e[k] - diff) <= 1e-9 * mag[k], (k, e[k], diff, mag[k])
        assert abs(diff) > 1e-3 * mag[k]  # the check is not vacuous


def test_assembled_operator_is_the_gradient_jacobian_in_tension():
    # assembleFineOperator (newtonMultigrid.ts:169-274) against a central-difference Jacobian of
    # nodeGradient (solver.ts:387-422). Axial springs and shear diagonals in tension with linear
    # laws, no bending and no air drag: there the clamp projections (max(K, 0), max(f, 0) / l) are
    # inactive and the projected Hessian is the exact one, so a factor error in any block (mass,
    # grip, axial, shear, viscosity) fails. Bending is left out because its block is a majoriser
    # (kb (Psi' + 2 Psi'' q) w w^T), not the Hessian.
    n = 8
    S = _linear_sheet(n, bend_k=0.0, drag=0.0)
    S.position[:, 0] *= 1.03  # stretch both families (x: u, z: v) by 3 %
    S.position[:, 2] *= 1.02
    S.mass[:] = 0.5  # inertia m / h^2 = 7200 N/m, of the order of the span stiffness
    grip = {"nodes": [0, 1], "weights": [0.5, 0.5], "offsets": np.zeros((2, 3)),
            "goal": S.position[[0, 1]].mean(0) + np.array([0, 1e-3, -1e-3]), "k": 2e4}
    o = O.Oracle(S, make_params(h=1 / 120, newton_iterations=1, history=False), [grip])
    o.predict()
    r = np.random.default_rng(11)
    free = np.repeat(o.pinned == 0, 3)
    # an iterate a fraction of a percent of the pitch away: every span and diagonal stays stretched
    o.dx[:] = np.where(free, 1e-6 * r.standard_normal(3 * S.node_count), 0.0)
    o.assemble()
    A = o.stencil(0)
    N = S.node_count
    eps = 1e-8
    worst = 0.0
    scale = np.max(np.abs(A))
    for m in range(N):
        if o.pinned[m]:
            continue
        im, jm = m % n, m // n
        nbrs = [q for q in range(N) if abs(q % n - im) <= 2 and abs(q // n - jm) <= 2 and not o.pinned[q]]
        for c in range(3):
            base = o.dx[3 * m + c]
            o.dx[3 * m + c] = base + eps
            gp = {q: o.node_gradient(q) for q in nbrs}
            o.dx[3 * m + c] = base - eps
            gm = {q: o.node_gradient(q) for q in nbrs}
            o.dx[3 * m + c] = base
            for q in nbrs:
                col = (gp[q] - gm[q]) / (2 * eps)
                off = ((m // n - q // n) + 2) * 5 + (m % n - q % n) + 2
                worst = max(worst, np.max(np.abs(A[q, off][:, c] - col)))
    tol = 1e-6 * scale
    assert worst <= tol, (worst, scale)
    # every term is far above the tolerance: inertia, the grip (k w = 1e4), the axial viscosity
    # c = visc / (L h) and the transverse tension stiffness (the smallest, f / l on y)
    visc = S.layers[0].material["yarns"][0]["axialViscosity"] / (1e-3 / 120)
    assert min(0.5 * 120 ** 2, 1e4, visc) > 100 * tol
    assert np.min(A[~o.pinned.astype(bool), 12][:, 1, 1] - 0.5 * 120 ** 2) > 100 * tol
    # the tension really holds: every span's tangent block is the unclamped one
    assert np.all(np.abs(A[~o.pinned.astype(bool), 12]).sum(axis=(1, 2)) > 0)

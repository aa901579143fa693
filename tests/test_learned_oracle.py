"""The FP64 restatement of the learned coarse-space network (oracle/learned_oracle.py) pinned by
checks that do not restate it: central finite differences of its loss, the loss's closed forms
(0 when the fields span the targets' errors, 1 when they cannot see the targets), and its level
operator apply against the FP64 solver oracle's Galerkin level."""
import numpy as np

import learned_oracle as LO
import oracle as O
from bd.capi import make_params
from bd.configs import sheet_scene
from bd.material import linear_sheet_material


def level_planes(stencil, nx, ny):
    """The oracle's (N, 25, 3, 3) level stencil -> the 78-plane half storage (centre and the 12
    forward blocks, symmetric entries xx xy xz yy yz zz)."""
    N = nx * ny
    op = np.zeros((78, N))
    for c in range(13):
        di, dj = LO.COARSE_DI[c], LO.COARSE_DJ[c]
        B = stencil[:, (dj + 2) * 5 + di + 2]
        for q, (a, b) in enumerate(((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))):
            op[6 * c + q] = B[:, a, b]
    return op


def small_level():
    mat = linear_sheet_material(1e-3, 2e4, 1.6, 1e-5, 0.35, 2e4 / 2.4)
    S = sheet_scene(33, 1e-3, mat)
    o = O.Oracle(S, make_params(max_coarsest_nodes=100))
    o.predict()
    o.assemble()
    o.build_hierarchy()
    nx = ny = 17
    st = o.stencil(1)
    return o, level_planes(st, nx, ny), o.fixed(1), nx, ny


def test_level_apply_matches_the_solver_oracle():
    o, op, fixed, nx, ny = small_level()
    x = np.random.default_rng(1).standard_normal((nx * ny, 3))
    x[fixed != 0] = 0
    y = LO.coarse_apply(op, x, nx, ny, fixed)
    ref = o.apply(1, x.ravel()).reshape(-1, 3)
    ref[fixed != 0] = 0
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))


def sample(seed=3):
    o, op, fixed, nx, ny = small_level()
    rng = np.random.default_rng(seed)
    e = [rng.standard_normal((nx * ny, 3)) for _ in range(2)]
    for v in e:
        v[fixed != 0] = 0
    b = [LO.coarse_apply(op, v, nx, ny, fixed) for v in e]
    s = [float(np.sum(v * bb)) for v, bb in zip(e, b)]
    return op, fixed, nx, ny, e, b, s


def test_loss_closed_forms():
    op, fixed, nx, ny, e, b, s = sample()
    rng = np.random.default_rng(4)
    z = rng.standard_normal((nx * ny, 24))
    z[:, 0:3] = e[0]
    z[:, 3:6] = e[1]
    z[fixed != 0] = 0
    loss, _, G = LO.loss_and_dz(z, op, nx, ny, fixed, b, s)
    assert abs(loss) < 1e-9  # both errors lie in span(Z)
    # fields that cannot see the targets: targets on the left third, fields on the right third
    # (beyond the radius-2 stencil), so c = Z^T b = 0 and nothing of the energy is captured
    i = np.arange(nx * ny) % nx
    bl = [np.where((i < nx // 3)[:, None], bb, 0.0) for bb in b]
    zr = np.where((i > 2 * nx // 3)[:, None], z, 0.0)
    loss, _, _ = LO.loss_and_dz(zr, op, nx, ny, fixed, bl, s)
    assert abs(loss - 1.0) < 1e-12
    # the Gram matrix is Z^T A Z with the apply (symmetric positive)
    assert np.allclose(G, G.T, rtol=1e-10, atol=0)
    assert np.all(np.linalg.eigvalsh(G) > 0)


def test_gradient_matches_finite_differences():
    op, fixed, nx, ny, e, b, s = sample()
    w = LO.init_weights(5)
    loss, g = LO.loss_grad(w, op, fixed, nx, ny, b, s)
    assert 0 < loss < 1
    rng = np.random.default_rng(6)
    checked = 0
    for li, (i, o, wo, bo) in enumerate(LO.LAYERS):
        idx = list(rng.integers(wo, wo + min(i, 17 if li == 0 else i) * o, 6)) + list(rng.integers(bo, bo + o, 2))
        if li == len(LO.LAYERS) - 1:
            idx = [k for k in idx if (k - wo) % o < 24 and k < bo] + [bo + 1, bo + 5]
        for k in idx:
            eps = 1e-6 * max(1.0, abs(w[k]))
            wp, wm = w.copy(), w.copy()
            wp[k] += eps
            wm[k] -= eps
            fd = (LO.loss_grad(wp, op, fixed, nx, ny, b, s)[0] - LO.loss_grad(wm, op, fixed, nx, ny, b, s)[0]) / (2 * eps)
            assert abs(fd - g[k]) <= 1e-5 * np.max(np.abs(g)) + 1e-4 * abs(g[k]), (li, k, fd, g[k])
            checked += 1
    assert checked >= 40
    assert np.max(np.abs(g)) > 0

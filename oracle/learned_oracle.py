"""FP64 restatement of the learned coarse-space network, its loss and its gradient.  TEST
INFRASTRUCTURE ONLY: only tests/ import it, as the checker of the GPU's forward and backward kernels
(break-that-denim_amd/csrc/learned.hip). The reference has no learned part; this module restates the
build's own network from its definition:

  features   per node of the net's level: u, v, sin/cos(pi u), sin/cos(pi v), sin/cos(2 pi u),
             sin/cos(2 pi v), fixed flag, asinh(diagonal block / mean diagonal), the restricted
             Newton right-hand side over its rms (20 of 32 inputs)
  encoder    h0 = relu(x We + be)                                      (32 -> 64)
  4 layers   h  = h + relu([h, mean of the lattice neighbours' h] W + b)  (128 -> 64)
  decoder    z  = h Wd + bd                                            (64 -> 32, 24 used: 8 fields x 3)
  loss       L = sum_t w_t (1 - c_t^T y_t / s_t),  c_t = Z^T b_t,  y_t = (Z^T A Z)^+ c_t

`emulate_bf16` rounds at the points where the GPU stores bf16 (inputs, layer outputs, neighbour
means, the weights; in the backward the output gradient and every ReLU-masked gradient), so the
GPU can be compared tightly; without it everything is FP64 (the form the finite-difference checks
pin).

Parity pinning (tests/test_learned_oracle.py): the analytic gradient against central finite
differences of the loss; the loss against its closed form on fields that span the targets exactly
(loss 0) and on fields orthogonal to them (loss 1); the Gram matrix against a dense Z^T A Z.
"""
from __future__ import annotations

import math

import numpy as np

N_IN, N_HID, N_MP, N_OUT, K_FIELDS = 32, 64, 4, 32, 8
PINV_TOL = 1e-7

# (in, out, weight offset, bias offset) of the flat parameter vector, weights in-major W[i][o]
LAYERS = [(N_IN, N_HID, 0, N_IN * N_HID)]
_off = N_IN * N_HID + N_HID
for _ in range(N_MP):
    LAYERS.append((2 * N_HID, N_HID, _off, _off + 2 * N_HID * N_HID))
    _off += 2 * N_HID * N_HID + N_HID
LAYERS.append((N_HID, N_OUT, _off, _off + N_HID * N_OUT))
N_PARAMS = _off + N_HID * N_OUT + N_OUT

# forward offsets of the coarse 13-block stencil (block 0 the centre)
COARSE_DI = [0, 1, 2, -2, -1, 0, 1, 2, -2, -1, 0, 1, 2]
COARSE_DJ = [0, 0, 0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2]


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def split(w):
    """Flat parameters -> [(W (in, out), b (out,))]."""
    return [(w[wo:wo + i * o].reshape(i, o), w[bo:bo + o]) for i, o, wo, bo in LAYERS]


def lattice(nx, ny):
    """Per node the up to 4 lattice neighbours (-1: none) and the degree (one layer, row-major)."""
    N = nx * ny
    n = np.arange(N)
    i, j = n % nx, n // nx
    nb = np.stack([np.where(i > 0, n - 1, -1), np.where(i + 1 < nx, n + 1, -1),
                   np.where(j > 0, n - nx, -1), np.where(j + 1 < ny, n + nx, -1)], 1)
    return nb, (nb >= 0).sum(1)


def features(op, fixed, nx, ny, b):
    """(N, 32) inputs from the level operator (78 planes, N used entries each), fixed flags and the
    restricted right-hand side b (N, 3)."""
    N = nx * ny
    n = np.arange(N)
    u = (n % nx) / (nx - 1)
    v = (n // nx) / (ny - 1)
    free = fixed[:N] == 0
    diag = (op[0, :N].astype(np.float64) + op[3, :N] + op[5, :N]) / 3.0
    scale = diag[free].mean() if free.any() else 1.0
    X = np.zeros((N, N_IN))
    cols = [u, v, np.sin(np.pi * u), np.sin(np.pi * v), np.cos(np.pi * u), np.cos(np.pi * v),
            np.sin(2 * np.pi * u), np.sin(2 * np.pi * v), np.cos(2 * np.pi * u), np.cos(2 * np.pi * v),
            (~free).astype(float)]
    for c, col in enumerate(cols):
        X[:, c] = col
    for q in range(6):
        X[:, 11 + q] = np.where(free, np.arcsinh(op[q, :N].astype(np.float64) / scale), 0.0)
    b = np.asarray(b, dtype=np.float64)[:N]
    rms = np.sqrt(np.sum(b[free] ** 2) / (3 * free.sum())) if free.any() and np.any(b[free]) else 1.0
    X[:, 17:20] = np.where(free[:, None], b / rms, 0.0)
    return X


def forward(w, X, nx, ny, fixed, emulate_bf16=False):
    """z (N, 24) and the activations the backward needs."""
    R = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    P = [(R(W), b) for W, b in split(w)]
    nb, deg = lattice(nx, ny)
    x = R(X)
    a0 = x @ P[0][0] + P[0][1]
    h = [R(np.maximum(a0, 0))]
    apre = [a0]
    means = []
    for l in range(1, N_MP + 1):
        hp = h[-1]
        s = np.zeros_like(hp)
        for q in range(4):
            m = nb[:, q] >= 0
            s[m] += hp[nb[m, q]]
        mean = R(s / np.maximum(deg, 1)[:, None])
        a = np.concatenate([hp, mean], 1) @ P[l][0] + P[l][1]
        h.append(R(hp + np.maximum(a, 0)))
        apre.append(a)
        means.append(mean)
    z = h[-1] @ P[-1][0] + P[-1][1]
    z = z[:, :3 * K_FIELDS]
    z[fixed[:len(z)] != 0] = 0
    return z, {"x": x, "h": h, "apre": apre, "means": means, "P": P, "nb": nb, "deg": deg}


def coarse_apply(op, x, nx, ny, fixed):
    """y = A x for the 13-block half-stored coarse operator (78 planes), x and y (N, 3); fixed rows 0."""
    N = nx * ny
    n = np.arange(N)
    i, j = n % nx, n // nx
    y = np.zeros((N, 3))

    def blk(c, rows):
        s = op[6 * c:6 * c + 6, rows].astype(np.float64).T
        B = np.empty((len(rows), 3, 3))
        B[:, 0, 0], B[:, 0, 1], B[:, 0, 2] = s[:, 0], s[:, 1], s[:, 2]
        B[:, 1, 0], B[:, 1, 1], B[:, 1, 2] = s[:, 1], s[:, 3], s[:, 4]
        B[:, 2, 0], B[:, 2, 1], B[:, 2, 2] = s[:, 2], s[:, 4], s[:, 5]
        return B

    y += np.einsum("nab,nb->na", blk(0, n), x)
    for c in range(1, 13):
        di, dj = COARSE_DI[c], COARSE_DJ[c]
        f = (i + di >= 0) & (i + di < nx) & (j + dj < ny)
        rows = n[f]
        y[rows] += np.einsum("nab,nb->na", blk(c, rows), x[rows + dj * nx + di])
        bk = (i - di >= 0) & (i - di < nx) & (j - dj >= 0)
        rows = n[bk]
        m = rows - dj * nx - di
        y[rows] += np.einsum("nab,nb->na", blk(c, m), x[m])
    y[fixed[:N] != 0] = 0
    return y


def pinv_gram(G):
    """Pseudo-inverse of the symmetric Gram matrix: eigenvalues below PINV_TOL times the largest are
    dropped (the GPU's Jacobi decomposition does the same)."""
    w, V = np.linalg.eigh(0.5 * (G + G.T))
    d = np.where((w > PINV_TOL * w.max()) & (w > 0), 1.0 / np.where(w > 0, w, 1.0), 0.0)
    return (V * d) @ V.T


def fields(z):
    """(N, 24) decoder output -> list of 8 (N, 3) fields (column 3 k + q is field k, component q)."""
    return [z[:, 3 * k:3 * k + 3] for k in range(K_FIELDS)]


def loss_and_dz(z, op, nx, ny, fixed, b, s):
    """Loss and dL/dz (N, 24) for targets b (list of (N, 3)) with energies s (targets with s <= 0 are
    left out)."""
    Z = fields(z)
    AZ = [coarse_apply(op, f, nx, ny, fixed) for f in Z]
    free = (fixed[:len(z)] == 0)[:, None]
    G = np.array([[np.sum(Z[a] * AZ[c] * free) for c in range(K_FIELDS)] for a in range(K_FIELDS)])
    Ginv = pinv_gram(G)
    on = [t for t in range(len(b)) if s[t] > 0]
    loss = 0.0
    dz = np.zeros_like(z)
    for t in on:
        c = np.array([np.sum(Z[a] * b[t] * free) for a in range(K_FIELDS)])
        y = Ginv @ c
        loss += (1 - c @ y / s[t]) / len(on)
        r = (b[t] - sum(AZ[a] * y[a] for a in range(K_FIELDS))) * free
        coef = -2.0 / (len(on) * s[t])
        for a in range(K_FIELDS):
            dz[:, 3 * a:3 * a + 3] += coef * r * y[a]
    return loss, dz, G


def backward(dz, cache, emulate_bf16=False):
    """Flat gradient of the parameters from dL/dz."""
    R = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    P, h, apre, means, nb, deg, x = (cache[k] for k in ("P", "h", "apre", "means", "nb", "deg", "x"))
    g = np.zeros(N_PARAMS)
    dzf = np.zeros((dz.shape[0], N_OUT))
    dzf[:, :dz.shape[1]] = dz
    dzf = R(dzf)

    def put(layer, dW, db):
        i, o, wo, bo = LAYERS[layer]
        g[wo:wo + i * o] = dW.ravel()
        g[bo:bo + o] = db

    put(N_MP + 1, h[-1].T @ dzf, dzf.sum(0))
    dh = dzf @ P[-1][0].T
    for l in range(N_MP, 0, -1):
        da = R(dh * (apre[l] > 0))
        U = np.concatenate([h[l - 1], means[l - 1]], 1)
        put(l, U.T @ da, da.sum(0))
        dU = da @ P[l][0].T
        dm = dU[:, N_HID:] / np.maximum(deg, 1)[:, None]
        back = np.zeros_like(dh)
        for q in range(4):
            m = nb[:, q] >= 0
            back[m] += dm[nb[m, q]]
        dh = dh + dU[:, :N_HID] + back
    da0 = R(dh * (apre[0] > 0))
    put(0, x.T @ da0, da0.sum(0))
    return g


def loss_grad(w, op, fixed, nx, ny, b, s, emulate_bf16=False):
    """Loss and flat gradient of one sample (operator (78, >=N), fixed (>=N,), targets list of (N, 3))."""
    X = features(op, fixed, nx, ny, b[0])
    z, cache = forward(w, X, nx, ny, fixed, emulate_bf16)
    loss, dz, _ = loss_and_dz(z, op, nx, ny, fixed, b, s)
    return loss, backward(dz, cache, emulate_bf16)


def init_weights(seed=0):
    """A random parameter vector (for tests): uniform Glorot, message-passing layers scaled by 0.3."""
    rng = np.random.default_rng(seed)
    w = np.zeros(N_PARAMS)
    for li, (i, o, wo, bo) in enumerate(LAYERS):
        a = math.sqrt(6.0 / (i + o)) * (0.3 if 1 <= li <= N_MP else 1.0)
        w[wo:wo + i * o] = rng.uniform(-a, a, i * o)
        w[bo:bo + o] = rng.uniform(-0.1, 0.1, o)
    return w

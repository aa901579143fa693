"""Scene compilation: woven lattices with mounts, wear, per-yarn strength and initial span state.

Restates the reference's scene compiler (src/simulation/scene.ts:48-428) as vectorised numpy:
node n = offset + j * nx + i (i = warp column along u, j = pick row along v); node n owns span
2n (to (i+1, j), u family) and span 2n + 1 (to (i, j+1), v family); tails per node and side
[u-, u+, v-, v+]. The yarn damage field holds per crossing c and family f (index 2c + f) a peak
tension, an initial damage code and an abraded cross-section code (src/simulation/yarnField.ts:26-31).

Output arrays are what the solver library consumes (bd.h `bd_scene`), float64 unless noted.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import json
import math

import numpy as np

from .material import DENIM, FAMILY_U, FAMILY_V, bundle_weave, crossing_mass
from .noise import fbm, noise1, rand01, smoothstep

SEG_EXISTS = 1
SEG_BROKEN = 2
SEG_INITIAL_DEFECT = 4
SEG_SLIPPED = 8

DAMAGE_CUT = 255
DAMAGE_PULLED = 254
DAMAGE_SEPARATED = 254
MAX_BUNDLE = 32
AREA_MIN = 0.02

DRAPE_SEED = 101          # scene.ts:15
MECHANICS_VERSION = 10    # scene.ts:131
FIELD_GRID = 1.0e-3       # scene.ts:346


def area_code(a):
    """u8 cross-section code of fraction a (yarnField.ts:77-79); JS Math.round = floor(x + 0.5)."""
    return np.clip(np.floor(np.asarray(a, dtype=np.float64) * 255 + 0.5), 1, 255).astype(np.uint8)


def area_of(code):
    return np.maximum(AREA_MIN, np.asarray(code, dtype=np.float64) / 255.0)


NO_WEAR = {"center": [0, 0], "axis": [1, 0], "radii": [1, 1], "amount": 0, "cellSize": 0,
           "featureScale": 0.01, "defectThreshold": 2, "defectDensity": 0, "sparseFrom": 2}


def layer_spec(material=DENIM, size_u=0.01, size_v=0.01, **kw):
    """A LayerSpec (scene.ts:48-71) as a dict with the reference's defaults for a plain coupon."""
    spec = {
        "material": material, "bundle": [1, 1], "sizeU": size_u, "sizeV": size_v, "z": 0.0,
        "angle": 0.0, "compressU": 0.0, "compressV": 0.0, "shear": 0.0, "collarRows": 1,
        "wear": dict(NO_WEAR), "wearStrengthLoss": [0.0, 0.0], "seed": 1, "perturbation": 0.0,
    }
    spec.update(kw)
    return spec


def coupon(nu, nv, material=DENIM, **kw):
    """Coupon of nu x nv nodes as the reference tests build it (newtonMultigrid.test.ts:27-43)."""
    bundle = kw.get("bundle", [1, 1])
    return layer_spec(material, size_u=(nu - 1) * material["pitchU"] * bundle[0],
                      size_v=(nv - 1) * material["pitchV"] * bundle[1], **kw)


@dataclass
class Layer:
    index: int
    material: dict       # mechanical (bundled) material
    weave: dict          # individual-yarn material
    bundle: tuple
    nx: int
    ny: int
    node_offset: int
    node_count: int
    mapA: tuple
    mapOrigin: tuple
    restCenter: tuple
    # yarn damage field layout (yarnField.ts:58-66)
    field_offset: int = 0
    K: int = 0
    M: int = 0

    @property
    def bu(self):
        return self.bundle[0]

    @property
    def bv(self):
        return self.bundle[1]


@dataclass
class CompiledScene:
    layers: list
    node_count: int
    rest_uv: np.ndarray
    position: np.ndarray          # (N, 3)
    mass: np.ndarray              # (N,)
    pinned: np.ndarray            # (N,) uint8
    wear: np.ndarray
    structural_wear: np.ndarray
    rest_length: np.ndarray       # (2N,)
    softening: np.ndarray         # (2N,)
    seg_flags: np.ndarray         # (2N,) uint16
    intact: np.ndarray            # (2N,)
    cut_mask: np.ndarray          # (2N,) uint32
    fail_site: np.ndarray         # (2N,) float32 in the reference; kept as float64 here
    crossing_count: int
    field_strength: np.ndarray    # (2C,) float32
    field_damage: np.ndarray      # (2C,) uint8
    field_area: np.ndarray        # (2C,) uint8
    crossing_normal0: np.ndarray  # (N,)
    tails: np.ndarray             # (4N,)
    extra: dict = field(default_factory=dict)

    def copy(self):
        import copy as _c
        return _c.deepcopy(self)


def _worley(seed, x, y):
    """Worley F1/F2 on a jittered unit grid (scene.ts:153-174), vectorised."""
    ix = np.floor(x).astype(np.int64)
    iy = np.floor(y).astype(np.int64)
    f1 = np.full(np.shape(x), 1e9)
    f2 = np.full(np.shape(x), 1e9)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            cx = ix + dx
            cy = iy + dy
            px = cx + 0.15 + 0.7 * rand01(seed, cx, cy)
            py = cy + 0.15 + 0.7 * rand01(seed + 17, cx, cy)
            d = np.hypot(px - x, py - y)
            lt1 = d < f1
            lt2 = (~lt1) & (d < f2)
            f2 = np.where(lt1, f1, np.where(lt2, d, f2))
            f1 = np.where(lt1, d, f1)
    return f1, f2


def wear_at(spec, seed, x, y):
    """[structural, visible] wear at rest-plane points (scene.ts:182-212), vectorised."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    if spec["amount"] <= 0:
        z = np.zeros(np.broadcast(x, y).shape)
        return z, z.copy()
    al = math.hypot(spec["axis"][0], spec["axis"][1]) or 1.0
    ax = spec["axis"][0] / al
    ay = spec["axis"][1] / al
    dx = x - spec["center"][0]
    dy = y - spec["center"][1]
    a = (dx * ax + dy * ay) / spec["radii"][0]
    b = (-dx * ay + dy * ax) / spec["radii"][1]
    S = spec["featureScale"]
    warp = 0.32 * fbm(seed + 3, x / (1.4 * S), y / (1.4 * S), 4)
    r = np.hypot(a, b) + warp
    w = 1 - smoothstep(0.4, 1.1, r)
    cells = np.ones_like(w)
    if spec["cellSize"] > 0:
        jx = x + 0.4 * S * fbm(seed + 11, x / S, y / S, 3)
        jy = y + 0.4 * S * fbm(seed + 12, x / S, y / S, 3)
        f1, f2 = _worley(seed + 5, jx / spec["cellSize"], jy / spec["cellSize"])
        cells = np.power(smoothstep(0.03, 0.3, f2 - f1), 0.7)
    w = w * (0.82 + 0.18 * fbm(seed + 7, x / (0.3 * S), y / (0.3 * S), 3)) * spec["amount"]
    structural = np.clip(w * (0.12 + 0.88 * cells), 0, 1)
    visible = np.clip(np.power(w, 2.2) * (0.35 + 0.65 * cells), 0, 1)
    return structural, visible


def compile_layer(spec, index, node_offset, nx=None, ny=None):
    """Layer layout and mount map (scene.ts:214-231). nx/ny may be given explicitly: the
    reference's floor(size / pitch) + 1 is rounding-fragile for sizes that are pitch multiples."""
    bundle = tuple(spec.get("bundle", [1, 1]))
    m = bundle_weave(spec["material"], bundle[0], bundle[1])
    if nx is None:
        nx = int(math.floor(spec["sizeU"] / m["pitchU"])) + 1
    if ny is None:
        ny = int(math.floor(spec["sizeV"] / m["pitchV"])) + 1
    c = math.cos(spec["angle"])
    s = math.sin(spec["angle"])
    k00 = 1 - spec["compressU"]
    k01 = spec["shear"]
    k11 = 1 - spec["compressV"]
    mapA = (c * k00, c * k01 - s * k11, s * k00, s * k01 + c * k11)
    restCenter = ((nx - 1) * m["pitchU"] / 2, (ny - 1) * m["pitchV"] / 2)
    return Layer(index, m, spec["material"], bundle, nx, ny, node_offset, nx * ny, mapA,
                 (0.0, 0.0, spec["z"]), restCenter)


def _rest_plane_point(spec, L, u, v):
    c = math.cos(spec["angle"])
    s = math.sin(spec["angle"])
    du = u - L.restCenter[0]
    dv = v - L.restCenter[1]
    return c * du - s * dv, s * du + c * dv


def _alloc(layers, N, C):
    return CompiledScene(
        layers=layers, node_count=N, rest_uv=np.zeros((N, 2)), position=np.zeros((N, 3)),
        mass=np.zeros(N), pinned=np.zeros(N, np.uint8), wear=np.zeros(N, np.float32),
        structural_wear=np.zeros(N, np.float32), rest_length=np.zeros(2 * N),
        softening=np.zeros(2 * N), seg_flags=np.zeros(2 * N, np.uint16), intact=np.zeros(2 * N),
        cut_mask=np.zeros(2 * N, np.uint32), fail_site=np.zeros(2 * N), crossing_count=C,
        field_strength=np.zeros(2 * C, np.float32), field_damage=np.zeros(2 * C, np.uint8),
        field_area=np.full(2 * C, 255, np.uint8), crossing_normal0=np.zeros(N), tails=np.zeros(4 * N))


def _fill_layer(S, spec, L):
    """Positions, mass, mounts, wear and span rest state of one layer (scene.ts:289-343)."""
    m = L.material
    nx, ny, off = L.nx, L.ny, L.node_offset
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    i = i.ravel()
    j = j.ravel()
    n = off + j * nx + i
    u = i * m["pitchU"]
    v = j * m["pitchV"]
    S.rest_uv[n, 0] = u
    S.rest_uv[n, 1] = v
    du = u - L.restCenter[0]
    dv = v - L.restCenter[1]
    A = L.mapA
    px = L.mapOrigin[0] + A[0] * du + A[1] * dv
    py = L.mapOrigin[1] + A[2] * du + A[3] * dv
    cr = spec["collarRows"]
    pinned = (i < cr) | (j < cr) | (i >= nx - cr) | (j >= ny - cr)
    fade_in = 1 - smoothstep(0.16, 0.24, np.hypot(px, py))
    pz = np.where(pinned, 0.0, spec["perturbation"] * fade_in * fbm(DRAPE_SEED, px / 0.04, py / 0.04, 3))
    S.position[n, 0] = px
    S.position[n, 1] = py
    S.position[n, 2] = L.mapOrigin[2] + pz
    S.mass[n] = crossing_mass(m)
    S.pinned[n] = pinned.astype(np.uint8)
    rx, ry = _rest_plane_point(spec, L, u, v)
    w, wvis = wear_at(spec["wear"], spec["seed"], rx, ry)
    S.wear[n] = wvis
    S.structural_wear[n] = w
    S.crossing_normal0[n] = m["crossingNormal0"] * (1 - 0.7 * w)
    for f in (0, 1):
        exists = (i < nx - 1) if f == FAMILY_U else (j < ny - 1)
        s = 2 * n[exists] + f
        y = m["yarns"][f]
        S.seg_flags[s] = SEG_EXISTS
        S.rest_length[s] = m["pitchU"] if f == FAMILY_U else m["pitchV"]
        S.intact[s] = 1.0
        S.softening[s] = y["softeningOpening"]
        S.fail_site[s] = 0.5


def _fill_field(S, spec, L):
    """Per-yarn peak tensions, cross-sections and initial cuts (scene.ts:355-428)."""
    if L.bu > MAX_BUNDLE or L.bv > MAX_BUNDLE:
        raise ValueError("bundle too large for the cut masks")
    weave = L.weave
    seed = spec["seed"]
    pu = weave["pitchU"]
    pv = weave["pitchV"]
    u0 = -L.bu * pu
    v0 = -L.bv * pv
    gw = int(math.ceil((L.K + 2 * L.bu) * pu / FIELD_GRID)) + 2
    gh = int(math.ceil((L.M + 2 * L.bv) * pv / FIELD_GRID)) + 2
    gy, gx = np.meshgrid(np.arange(gh), np.arange(gw), indexing="ij")
    rx, ry = _rest_plane_point(spec, L, u0 + gx * FIELD_GRID, v0 + gy * FIELD_GRID)
    g_wear = wear_at(spec["wear"], seed, rx, ry)[0].astype(np.float32).ravel()
    g_corr = fbm(seed + 503, rx / 0.011, ry / 0.011, 3).astype(np.float32).ravel()
    g_edge = fbm(seed + 991, rx / 0.004, ry / 0.004, 2).astype(np.float32).ravel()

    def sample(g, u, v):
        fx = np.minimum(gw - 1.001, np.maximum(0, (u - u0) / FIELD_GRID))
        fy = np.minimum(gh - 1.001, np.maximum(0, (v - v0) / FIELD_GRID))
        x = np.floor(fx).astype(np.int64)
        y = np.floor(fy).astype(np.int64)
        tx = fx - x
        ty = fy - y
        g = g.astype(np.float64)
        a = g[y * gw + x] + (g[y * gw + x + 1] - g[y * gw + x]) * tx
        b = g[(y + 1) * gw + x] + (g[(y + 1) * gw + x + 1] - g[(y + 1) * gw + x]) * tx
        return a + (b - a) * ty

    thr = spec["wear"]["defectThreshold"]
    mm, k = np.meshgrid(np.arange(L.M), np.arange(L.K), indexing="ij")
    mm = mm.ravel()
    k = k.ravel()
    v = (mm - (L.bv - 1) / 2) * pv
    u = (k - (L.bu - 1) / 2) * pu
    c = L.field_offset + mm * L.K + k
    w = sample(g_wear, u, v)
    corr = sample(g_corr, u, v)
    for f in (0, 1):
        loss = spec["wearStrengthLoss"][f]
        wf = np.power(w, 1.3) if f == FAMILY_U else np.power(w, 1.25)
        yarn_index = k if f == FAMILY_V else mm
        along = v if f == FAMILY_V else u
        slub = 1 + 0.25 * noise1(seed + 4099 + 13 * yarn_index + 7 * f, along / 0.0025)
        z = (rand01(seed + 503, c, f) + rand01(seed + 617, c, f) - 1) * 2.45
        scatter = np.exp((0.2 + 0.12 * w) * corr + (0.06 + 0.22 * w) * z)
        area = np.maximum(0.02, 1 - loss * wf)
        S.field_area[2 * c + f] = area_code(area)
        S.field_strength[2 * c + f] = (weave["yarns"][f]["strength"] * np.maximum(0.005, area * scatter * slub)).astype(np.float32)
    edge = thr + 0.04 * sample(g_edge, u, v) + 0.012 * (rand01(seed + 990, k) - 0.5)
    cut = (w > edge) & (rand01(seed + 991, k, mm) < spec["wear"]["defectDensity"])
    S.field_damage[2 * c[cut] + FAMILY_V] = DAMAGE_CUT
    R = weave["twillRepeat"]
    flt = np.floor((mm - (k % R) + 4 * R) / R).astype(np.int64)
    pf = 0.02 * np.power(smoothstep(spec["wear"]["sparseFrom"], thr, w), 2)
    on_face = ((mm - k) % R + R) % R < weave["twillWarpOver"]
    cut2 = (pf > 0) & on_face & (rand01(seed + 997, k, flt) < pf)
    S.field_damage[2 * c[cut2] + FAMILY_V] = DAMAGE_CUT


def field_layout(layers):
    """Crossing layout per layer: K = nx bu warps, M = ny bv picks (yarnField.ts:58-66)."""
    off = 0
    for L in layers:
        L.field_offset = off
        L.K = L.nx * L.bu
        L.M = L.ny * L.bv
        off += L.K * L.M
    return off


def span_crossings(L, f, i, j):
    """Crossing indices of span (i, j, f): array (..., B, P) for yarn q, p-th crossing (yarnField.ts:201-204)."""
    i = np.asarray(i)[..., None, None]
    j = np.asarray(j)[..., None, None]
    if f == 1:
        q = np.arange(L.bu)[:, None]
        p = np.arange(L.bv)[None, :]
        return L.field_offset + (j * L.bv + L.bv // 2 + p) * L.K + i * L.bu + q
    q = np.arange(L.bv)[:, None]
    p = np.arange(L.bu)[None, :]
    return L.field_offset + (j * L.bv + q) * L.K + i * L.bu + L.bu // 2 + p


def _part_span_yarns(L, damage, i, j, f, slip_end=-1):
    """Separated span's yarns leave the other side's weave (yarnField.ts:422-434)."""
    cr = span_crossings(L, f, i, j)
    B, P = cr.shape[-2:]
    half = P - P // 2
    for q in range(B):
        of_a = slip_end != 0
        for p in range(P):
            e = 2 * int(cr[q, p]) + f
            if damage[e] >= DAMAGE_SEPARATED:
                if slip_end < 0:
                    of_a = False
            elif of_a != (p < half):
                damage[e] = DAMAGE_PULLED


def couple_spans_initial(S):
    """Span state from the initial field (scene.ts:271-276 via yarnField.ts:443-499).

    Vectorised over spans with the per-yarn and per-crossing sums kept in the reference's
    sequential order; ruptures (rare, authored defects) are applied per span afterwards.
    """
    ruptures = 0
    before = S.seg_flags.copy()
    for L in S.layers:
        for f in (0, 1):
            jj, ii = np.meshgrid(np.arange(L.ny), np.arange(L.nx), indexing="ij")
            ok = (ii < L.nx - 1) if f == 0 else (jj < L.ny - 1)
            ii = ii[ok]
            jj = jj[ok]
            n = L.node_offset + jj * L.nx + ii
            s = 2 * n + f
            act = (S.seg_flags[s] & (SEG_EXISTS | SEG_BROKEN)) == SEG_EXISTS
            ii, jj, n, s = ii[act], jj[act], n[act], s[act]
            if len(s) == 0:
                continue
            cr = span_crossings(L, f, ii, jj)        # (ns, B, P)
            B, P = cr.shape[1:]
            dmg = S.field_damage[2 * cr + f].astype(np.int64)
            ar = area_of(S.field_area[2 * cr + f])
            total = np.zeros(len(s))
            mask = np.zeros(len(s), np.uint64)
            site = np.zeros(len(s))
            for q in range(B):
                dmax = np.zeros(len(s), np.int64)
                first = np.full(len(s), -1)
                comp = np.zeros(len(s))
                for p in range(P):
                    d = dmg[:, q, p]
                    dmax = np.maximum(dmax, d)
                    first = np.where((d >= DAMAGE_SEPARATED) & (first < 0), p, first)
                    comp = comp + 1 / ar[:, q, p]
                sep = dmax >= DAMAGE_SEPARATED
                total = total + np.where(sep, 0.0, 1 - dmax / DAMAGE_CUT) * (P / comp)
                mask = mask | np.where(sep, np.uint64(1) << np.uint64(q), np.uint64(0))
                site = site + np.where(sep, (first + 0.5) / P, 0.0)
            S.intact[s] = total / B
            S.cut_mask[s] = mask.astype(np.uint32)
            full = (1 << B) - 1
            for idx in np.nonzero(mask == np.uint64(full))[0]:
                sj = int(s[idx])
                nn = int(n[idx])
                phi = site[idx] / B
                rest = S.rest_length[sj]
                b = nn + 1 if f == 0 else nn + L.nx
                S.seg_flags[sj] |= SEG_BROKEN
                S.fail_site[sj] = phi
                S.tails[4 * nn + 2 * f + 1] += rest * phi
                S.tails[4 * b + 2 * f] += rest * (1 - phi)
                ruptures += 1
                _part_span_yarns(L, S.field_damage, int(ii[idx]), int(jj[idx]), f)
    newly = (S.seg_flags & SEG_BROKEN).astype(bool) & ~(before & SEG_BROKEN).astype(bool)
    S.seg_flags[newly] |= SEG_INITIAL_DEFECT
    return ruptures


def default_scene_specs():
    """defaultSceneSpecs (src/simulation/scene.ts:438-469): the product's 0.62 x 0.4 m worn denim
    coupon, 9 x 6 bundles, collar rows, seed 1001 (182 x 115 = 20,930 nodes)."""
    return [layer_spec(
        DENIM, bundle=[9, 6], sizeU=0.62, sizeV=0.4, z=0, angle=0.06, compressU=0.025, compressV=-0.002,
        shear=0.06, collarRows=2,
        wear=dict(center=[0.01, -0.006], axis=[1, -1.15], radii=[0.2, 0.14], amount=1, cellSize=0.075,
                  featureScale=0.05, defectThreshold=0.89, defectDensity=1, sparseFrom=0.6),
        wearStrengthLoss=[0.45, 0.9], seed=1001, perturbation=2.5e-3)]


def compile_scene(specs, sizes=None):
    """compileScene (scene.ts:233-278). `sizes` optionally fixes (nx, ny) per layer."""
    layers = []
    off = 0
    for l, spec in enumerate(specs):
        nx, ny = (sizes[l] if sizes else (None, None))
        L = compile_layer(spec, l, off, nx, ny)
        layers.append(L)
        off += L.node_count
    C = field_layout(layers)
    S = _alloc(layers, off, C)
    for spec, L in zip(specs, layers):
        _fill_layer(S, spec, L)
        _fill_field(S, spec, L)
    couple_spans_initial(S)
    return S


def scene_hash(S) -> int:
    """Content hash of the compiled initial state (scene.ts:134-150): FNV-1a style over the
    mechanics version, the materials' JSON, sampled field codes, f32 positions and span flags."""
    h = 0x811C9DC5

    def mix(v):
        nonlocal h
        h = ((h ^ (v & 0xFFFFFFFF)) * 0x01000193) & 0xFFFFFFFF

    mix(MECHANICS_VERSION)
    for L in S.layers:
        js = _js_json(L.material)
        for ch in js:
            mix(ord(ch))
    fd = S.field_damage
    fa = S.field_area
    for i in range(0, len(fd), 64):
        mix(int(fd[i]) ^ (int(fa[i]) << 16) ^ ((i << 8) & 0xFFFFFFFF))
    u = S.position.astype(np.float32).ravel().view(np.uint32)
    for x in u.tolist():
        mix(x)
    for x in S.seg_flags.tolist():
        mix(int(x))
    return h


def _js_number(x):
    """Number formatting of JSON.stringify (shortest round-trip, JS exponent style)."""
    if isinstance(x, bool):
        return "true" if x else "false"
    if isinstance(x, int) or (isinstance(x, float) and x.is_integer() and abs(x) < 1e21):
        return str(int(x))
    r = repr(float(x))
    if "e" in r:
        mant, exp = r.split("e")
        e = int(exp)
        if -7 < e < 21:
            return np.format_float_positional(float(x), trim="-")
        return f"{mant}e{'+' if e > 0 else '-'}{abs(e)}"
    return r


def _js_json(v):
    if isinstance(v, dict):
        return "{" + ",".join(json.dumps(k) + ":" + _js_json(x) for k, x in v.items()) + "}"
    if isinstance(v, (list, tuple)):
        return "[" + ",".join(_js_json(x) for x in v) + "]"
    if isinstance(v, str):
        return json.dumps(v)
    return _js_number(v)

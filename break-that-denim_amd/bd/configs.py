"""Workload scenes: hanging sheets with pinned top rows and moving grips.

The sheet configs express their physics through the reference's lattice model (the reference has no
StVK triangle membrane, Poisson coupling or dihedral bending): per-family linear axial laws with
E * pitch per span (warp : weft stiffness 1.6 : 1), bending rigidity k * pitch per span, lumped
mass areal_density * pitch^2 per crossing, and diagonal shear springs equal to the membrane shear
modulus E / (2 (1 + nu)) (a square cell's diagonal pair under simple shear stores
1/2 ks gamma^2 pitch^2). A breakable central region draws one break strain per yarn from
lognormal(ln median, sigma) with a fixed seed; its peak tension is the yarn law at that strain.
"""
from __future__ import annotations

import math

import numpy as np

from .capi import make_params
from .material import crossing_mass, linear_sheet_material
from .noise import fbm
from .scene import CompiledScene, Layer, SEG_EXISTS, field_layout

UNBREAKABLE = 1.0e9


def _axial_tension(y, eps):
    eps = np.asarray(eps, dtype=np.float64)
    ec = y["crimpStrain"]
    dk = y["eaHigh"] - y["eaLow"]
    lo = y["eaLow"] * eps + dk * eps * eps / (2 * ec)
    hi = y["eaLow"] * ec + dk * ec / 2 + y["eaHigh"] * (eps - ec)
    return np.where(eps <= 0, y["eaCompress"] * eps, np.where(eps <= ec, lo, hi))


def sheet_scene(n, pitch, material, perturbation=0.5e-3, seed=101, pin_top=True, yarn_region=0,
                break_median=0.12, break_sigma=0.3, break_seed=7):
    """n x n hanging sheet in the x-z plane (u -> x, v -> z, top row j = n-1 pinned), with a small
    out-of-plane perturbation along y from fbm noise to seed buckling."""
    nx = ny = n
    N = nx * ny
    L = Layer(0, material, material, (1, 1), nx, ny, 0, N, (1.0, 0.0, 0.0, 1.0), (0.0, 0.0, 0.0),
              ((nx - 1) * pitch / 2, (ny - 1) * pitch / 2))
    C = field_layout([L])
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    i = i.ravel()
    j = j.ravel()
    u = i * pitch - L.restCenter[0]
    v = j * pitch - L.restCenter[1]
    pinned = (j == ny - 1) if pin_top else np.zeros(N, bool)
    py = np.where(pinned, 0.0, perturbation * fbm(seed, u / 0.04, v / 0.04, 3))
    S = CompiledScene(
        layers=[L], node_count=N, rest_uv=np.stack([i * pitch, j * pitch], 1), position=np.stack([u, py, v], 1),
        mass=np.full(N, crossing_mass(material)), pinned=pinned.astype(np.uint8), wear=np.zeros(N, np.float32),
        structural_wear=np.zeros(N, np.float32), rest_length=np.zeros(2 * N), softening=np.zeros(2 * N),
        seg_flags=np.zeros(2 * N, np.uint16), intact=np.zeros(2 * N), cut_mask=np.zeros(2 * N, np.uint32),
        fail_site=np.zeros(2 * N), crossing_count=C, field_strength=np.full(2 * C, UNBREAKABLE, np.float32),
        field_damage=np.zeros(2 * C, np.uint8), field_area=np.full(2 * C, 255, np.uint8),
        crossing_normal0=np.full(N, material["crossingNormal0"]), tails=np.zeros(4 * N))
    n_idx = np.arange(N)
    for f in (0, 1):
        ex = (i < nx - 1) if f == 0 else (j < ny - 1)
        s = 2 * n_idx[ex] + f
        S.seg_flags[s] = SEG_EXISTS
        S.rest_length[s] = pitch
        S.intact[s] = 1.0
        S.softening[s] = material["yarns"][f]["softeningOpening"]
        S.fail_site[s] = 0.5
    if yarn_region > 0:
        rng = np.random.default_rng(break_seed)
        c0 = (n - yarn_region) // 2
        idx = np.arange(c0, c0 + yarn_region)
        # one break strain per yarn: warps (v family) per column, wefts (u family) per row
        eb_warp = np.exp(math.log(break_median) + break_sigma * rng.standard_normal(yarn_region))
        eb_weft = np.exp(math.log(break_median) + break_sigma * rng.standard_normal(yarn_region))
        kk, mm = np.meshgrid(idx, idx, indexing="xy")   # kk: column (warp), mm: row (weft)
        c = (mm * nx + kk).ravel()
        t_warp = _axial_tension(material["yarns"][1], eb_warp)[np.searchsorted(idx, kk.ravel())]
        t_weft = _axial_tension(material["yarns"][0], eb_weft)[np.searchsorted(idx, mm.ravel())]
        S.field_strength[2 * c + 1] = t_warp.astype(np.float32)
        S.field_strength[2 * c + 0] = t_weft.astype(np.float32)
        S.extra["yarn_region"] = (c0, yarn_region)
    return S


def corner_grip(S, corner_i, corner_j, k, size=3):
    """A size x size patch grip at a lattice corner: equal weights summing to 1, goal at the patch
    centroid, offsets from it (the ReferenceGrip quadratic spring, solver.ts:25-37)."""
    L = S.layers[0]
    i0 = 0 if corner_i == 0 else L.nx - size
    j0 = 0 if corner_j == 0 else L.ny - size
    nodes = [(j0 + b) * L.nx + (i0 + a) for b in range(size) for a in range(size)]
    pos = S.position[nodes]
    goal = pos.mean(0)
    return {"nodes": nodes, "weights": [1.0 / len(nodes)] * len(nodes), "offsets": pos - goal, "goal": goal.copy(), "k": k}


def config1():
    """64 x 64 sheet, 0.5 m, warp/weft 1.6:1 with E_warp = 2e4 N/m, nu = 0.2 (shear fit), bending
    k = 1e-5 N m, 0.35 kg/m^2, top row pinned, one bottom corner gripped and moved +x 2 mm/step,
    h = 1/120, 3 Newton iterations, 3-level MG (coarsest <= 400 nodes), rel tol 1e-5, no history."""
    n = 64
    pitch = 0.5 / 63
    mat = linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / (2 * 1.2))
    S = sheet_scene(n, pitch, mat)
    grip = corner_grip(S, 1, 0, k=2e4)
    params = make_params(h=1 / 120, newton_iterations=3, pcg_rel_tol=1e-5, pcg_max_iterations=100,
                         max_coarsest_nodes=400, coarse_sweeps=2, lag_coarse=True, line_search=True, history=False)
    return {"name": "sheet64-plain", "scene": S, "params": params, "grips": [grip], "velocity": [np.array([2e-3, 0, 0])],
            "steps": 120}


def config_sheet(n=1024, pitch=1e-3, region=256, rel_tol=1e-6, history=True, newton=3, self_contact=False):
    """n x n sheet at 1 mm pitch with a central breakable woven region (per-yarn lognormal break
    strains, median 0.12, sigma 0.3, seed 7) and two bottom-corner grips pulled apart at 5 mm/step;
    plain FP32 MG-PCG to rel-res `rel_tol`, with node-face self-contact when asked for."""
    mat = linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / (2 * 1.2))
    S = sheet_scene(n, pitch, mat, yarn_region=region)
    g0 = corner_grip(S, 0, 0, k=2e4)
    g1 = corner_grip(S, 1, 0, k=2e4)
    params = make_params(h=1 / 120, newton_iterations=newton, pcg_rel_tol=rel_tol, pcg_max_iterations=600,
                         max_coarsest_nodes=400, coarse_sweeps=2, lag_coarse=True, line_search=True, history=history,
                         self_contact=self_contact)
    return {"name": f"sheet{n}-yarns{region}" + ("-contact" if self_contact else ""), "scene": S, "params": params, "grips": [g0, g1],
            "velocity": [np.array([-5e-3, 0, 0]), np.array([5e-3, 0, 0])], "steps": 240}


def config_patches(count=48, side_range=(32, 192), speed_range=(1e-3, 8e-3), seed=4, pitch=1e-3, rel_tol=1e-6,
                   newton=3, steps=120):
    """`count` independent hanging patches (batched uneven solves): grid side uniform in `side_range`
    from `seed`, 1 mm pitch, the sheet material of the 1024 x 1024 workload, top row pinned, one bottom
    corner gripped and moved along a random in-plane direction at a speed uniform in `speed_range`
    (m per step). Each patch is its own Newton / MG-PCG system with the mixed method (hierarchy rebuilt
    every Newton iteration, coarsest level of at most 400 nodes solved exactly in FP64), rel-res
    `rel_tol`, h = 1/120."""
    rng = np.random.default_rng(seed)
    mat = linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / (2 * 1.2))
    patches = []
    for p in range(count):
        n = int(rng.integers(side_range[0], side_range[1] + 1))
        S = sheet_scene(n, pitch, mat, seed=101 + p)
        grip = corner_grip(S, int(rng.integers(0, 2)), 0, k=2e4)
        ang = rng.uniform(0.0, 2.0 * math.pi)
        speed = rng.uniform(*speed_range)
        params = make_params(h=1 / 120, newton_iterations=newton, pcg_rel_tol=rel_tol, pcg_max_iterations=200,
                             max_coarsest_nodes=400, coarse_sweeps=2, lag_coarse=False, line_search=True, history=False,
                             coarse_solve=1)
        patches.append({"side": n, "scene": S, "params": params, "grips": [grip],
                        "velocity": [np.array([speed * math.cos(ang), 0.0, speed * math.sin(ang)])]})
    return {"name": f"patches{count}", "patches": patches, "steps": steps}


def advance_grips(grips, velocity):
    """Move every grip goal by its per-step velocity (the caller's handle motion)."""
    for g, v in zip(grips, velocity):
        g["goal"] = np.asarray(g["goal"], dtype=np.float64) + v
    return grips


def fold_scene(nx=48, ny=64, pitch=1e-3, radius=1e-3, material=None):
    """A horizontal nx x ny sheet folded over itself along v: the lower layer at z = 0, a semicircular
    fold of the given radius, the upper layer at z = 2 radius (the folded-coupon workload's shape).
    Nodes follow the fold by arc length, so spans keep their rest length up to the chord of the arc."""
    mat = material or linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / (2 * 1.2))
    N = nx * ny
    L = Layer(0, mat, mat, (1, 1), nx, ny, 0, N, (1.0, 0.0, 0.0, 1.0), (0.0, 0.0, 0.0),
              ((nx - 1) * pitch / 2, (ny - 1) * pitch / 2))
    C = field_layout([L])
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    i = i.ravel()
    j = j.ravel()
    s = j * pitch
    arc = math.pi * radius
    L1 = ((ny - 1) * pitch - arc) / 2
    th = np.clip((s - L1) / radius, 0.0, math.pi)
    y = np.where(s <= L1, s, np.where(s >= L1 + arc, L1 - (s - L1 - arc), L1 + radius * np.sin(th)))
    z = np.where(s <= L1, 0.0, np.where(s >= L1 + arc, 2 * radius, radius * (1 - np.cos(th))))
    x = i * pitch - (nx - 1) * pitch / 2
    S = CompiledScene(
        layers=[L], node_count=N, rest_uv=np.stack([i * pitch, j * pitch], 1), position=np.stack([x, y - L1 / 2, z], 1),
        mass=np.full(N, crossing_mass(mat)), pinned=np.zeros(N, np.uint8), wear=np.zeros(N, np.float32),
        structural_wear=np.zeros(N, np.float32), rest_length=np.zeros(2 * N), softening=np.zeros(2 * N),
        seg_flags=np.zeros(2 * N, np.uint16), intact=np.zeros(2 * N), cut_mask=np.zeros(2 * N, np.uint32),
        fail_site=np.zeros(2 * N), crossing_count=C, field_strength=np.full(2 * C, UNBREAKABLE, np.float32),
        field_damage=np.zeros(2 * C, np.uint8), field_area=np.full(2 * C, 255, np.uint8),
        crossing_normal0=np.full(N, mat["crossingNormal0"]), tails=np.zeros(4 * N))
    n_idx = np.arange(N)
    for f in (0, 1):
        ex = (i < nx - 1) if f == 0 else (j < ny - 1)
        sg = 2 * n_idx[ex] + f
        S.seg_flags[sg] = SEG_EXISTS
        S.rest_length[sg] = pitch
        S.intact[sg] = 1.0
        S.softening[sg] = mat["yarns"][f]["softeningOpening"]
        S.fail_site[sg] = 0.5
    S.extra["fold"] = {"L1": L1, "radius": radius, "j_lower": int(np.floor(L1 / pitch)),
                       "j_upper": int(np.ceil((L1 + arc) / pitch))}
    return S


def yarn_region_strengths(S, region, material, break_median=0.12, break_sigma=0.3, break_seed=7, i0=None, j0=None):
    """Per-yarn lognormal break strains in a region x region block of crossings (lower-left corner
    (i0, j0), centred by default): one strain per warp (column) and per weft (row), turned into the
    peak tension of the yarn law at that strain; everything else stays unbreakable."""
    L = S.layers[0]
    nx, ny = L.nx, L.ny
    i0 = (nx - region) // 2 if i0 is None else i0
    j0 = (ny - region) // 2 if j0 is None else j0
    rng = np.random.default_rng(break_seed)
    eb_warp = np.exp(math.log(break_median) + break_sigma * rng.standard_normal(region))
    eb_weft = np.exp(math.log(break_median) + break_sigma * rng.standard_normal(region))
    kk, mm = np.meshgrid(np.arange(i0, i0 + region), np.arange(j0, j0 + region), indexing="xy")
    c = (mm * nx + kk).ravel()
    S.field_strength[2 * c + 1] = _axial_tension(material["yarns"][1], eb_warp)[(kk - i0).ravel()].astype(np.float32)
    S.field_strength[2 * c + 0] = _axial_tension(material["yarns"][0], eb_weft)[(mm - j0).ravel()].astype(np.float32)
    S.extra["yarn_region"] = (i0, j0, region)
    return S


def config2(n=256, region=64, pitch=1e-3, steps=240, rel_tol=1e-6, mixed=False):
    """The plain method at scale: an n x n sheet at 1 mm pitch folded over itself along v (the upper
    half turned over 2 mm above the lower one) and dropped under gravity with node-face self-contact
    (1 mm contact distance); a central region x region block of crossings, across the fold, breaks at
    per-yarn lognormal strains (median 0.12, sigma 0.3, seed 7), the rest is unbreakable; two 3 x 3
    grips at the lower layer's two front corners pulled apart at 5 mm/step (their goals led at most
    6 mm ahead of the held patch, the reference's 120 N grip force bound at k = 2e4 N/m);
    FP32 MG-PCG to rel-res 1e-6, history on. The Galerkin hierarchy is rebuilt at every Newton
    iteration: the contact terms change between iterations, and with the hierarchy of the first one
    the later solves need about 3x the iterations (and do not all reach 1e-6 within 400).
    mixed=True gives the mixed-precision method on the same scene: exact FP64 solve of a coarsest
    level of <= 1100 nodes, FP32 operators (FP16 coarse-level storage loses the contact terms' range:
    5x the iterations), FP64 reductions and residual replacement."""
    from .capi import BD_COARSE_EXACT
    mat = linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / (2 * 1.2))
    S = fold_scene(n, n, pitch, radius=1e-3, material=mat)
    yarn_region_strengths(S, region, mat)
    g0 = corner_grip(S, 0, 0, k=2e4)
    g1 = corner_grip(S, 1, 0, k=2e4)
    params = make_params(h=1 / 120, newton_iterations=3, pcg_rel_tol=rel_tol, pcg_max_iterations=400,
                         max_coarsest_nodes=1100 if mixed else 400, coarse_sweeps=2, lag_coarse=False, line_search=True,
                         history=True, self_contact=True, coarse_solve=BD_COARSE_EXACT if mixed else 0)
    return {"name": f"fold{n}-yarns{region}" + ("-mixed" if mixed else ""), "scene": S, "params": params,
            "grips": [g0, g1], "velocity": [np.array([-5e-3, 0, 0]), np.array([5e-3, 0, 0])], "steps": steps,
            "max_lead": 6e-3}


def led_goals(goals, velocity, centroids, max_lead):
    """Grip goals moved by their per-step velocity, each held within max_lead of its patch centroid
    (the reference's handle lead clamp, a bound on the grip force)."""
    out = []
    for g, v, c in zip(goals, velocity, centroids):
        t = np.asarray(g, dtype=np.float64) + v
        d = t - c
        n = np.linalg.norm(d)
        out.append(c + d * (max_lead / n) if n > max_lead else t)
    return np.array(out)


def config_fold(nx=48, ny=64, pitch=1e-3, radius=1e-3, floor_gap=0.05e-3, newton=3, rel_tol=1e-5):
    """The fold dropped under gravity onto a floor just below the lower layer, with self-contact:
    the upper layer lands on the lower one."""
    S = fold_scene(nx, ny, pitch, radius)
    half = 0.5 * S.layers[0].material["thickness"]
    params = make_params(h=1 / 120, newton_iterations=newton, pcg_rel_tol=rel_tol, pcg_max_iterations=100,
                         max_coarsest_nodes=400, coarse_sweeps=2, lag_coarse=True, line_search=True, history=False,
                         self_contact=True)
    return {"name": f"fold{nx}x{ny}", "scene": S, "params": params, "grips": [], "floor_z": -(half + floor_gap),
            "steps": 24}


def flap_scene(nx=48, ny=64, pitch=1e-3, gap=2e-3, material=None):
    """One horizontal lattice torn across its middle row into two flaps, the upper flap (rows
    j >= ny/2) turned over onto the lower one: reflected in v, shifted by half a pitch in both
    directions and lifted by `gap`. The spans across the tear are broken (SEG_BROKEN)."""
    mat = material or linear_sheet_material(pitch, 2e4, 1.6, 1e-5, 0.35, 2e4 / (2 * 1.2))
    N = nx * ny
    L = Layer(0, mat, mat, (1, 1), nx, ny, 0, N, (1.0, 0.0, 0.0, 1.0), (0.0, 0.0, 0.0),
              ((nx - 1) * pitch / 2, (ny - 1) * pitch / 2))
    C = field_layout([L])
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    i = i.ravel()
    j = j.ravel()
    half = ny // 2
    up = j >= half
    x = np.where(up, (i + 0.5) * pitch, i * pitch) - (nx - 1) * pitch / 2
    y = np.where(up, (ny - 1 - j + 0.5) * pitch, j * pitch) - (half - 1) * pitch / 2
    z = np.where(up, gap, 0.0)
    S = CompiledScene(
        layers=[L], node_count=N, rest_uv=np.stack([i * pitch, j * pitch], 1), position=np.stack([x, y, z], 1),
        mass=np.full(N, crossing_mass(mat)), pinned=np.zeros(N, np.uint8), wear=np.zeros(N, np.float32),
        structural_wear=np.zeros(N, np.float32), rest_length=np.zeros(2 * N), softening=np.zeros(2 * N),
        seg_flags=np.zeros(2 * N, np.uint16), intact=np.zeros(2 * N), cut_mask=np.zeros(2 * N, np.uint32),
        fail_site=np.zeros(2 * N), crossing_count=C, field_strength=np.full(2 * C, UNBREAKABLE, np.float32),
        field_damage=np.zeros(2 * C, np.uint8), field_area=np.full(2 * C, 255, np.uint8),
        crossing_normal0=np.full(N, mat["crossingNormal0"]), tails=np.zeros(4 * N))
    n_idx = np.arange(N)
    for f in (0, 1):
        ex = (i < nx - 1) if f == 0 else (j < ny - 1)
        sg = 2 * n_idx[ex] + f
        S.seg_flags[sg] = SEG_EXISTS
        S.rest_length[sg] = pitch
        S.intact[sg] = 1.0
        S.softening[sg] = mat["yarns"][f]["softeningOpening"]
        S.fail_site[sg] = 0.5
    tear = 2 * n_idx[j == half - 1] + 1
    S.seg_flags[tear] |= 2
    S.intact[tear] = 0.0
    S.extra["flaps"] = {"half": half, "gap": gap}
    return S


def config_flaps(nx=48, ny=64, pitch=1e-3, gap=2e-3, floor_gap=0.05e-3, newton=3, rel_tol=1e-5):
    """The upper flap falls under gravity onto the lower one, which lies on a floor just below it,
    with self-contact (the workload of a turned-over flap landing on the cloth)."""
    S = flap_scene(nx, ny, pitch, gap)
    half = 0.5 * S.layers[0].material["thickness"]
    params = make_params(h=1 / 120, newton_iterations=newton, pcg_rel_tol=rel_tol, pcg_max_iterations=100,
                         max_coarsest_nodes=400, coarse_sweeps=2, lag_coarse=True, line_search=True, history=False,
                         self_contact=True)
    return {"name": f"flaps{nx}x{ny}", "scene": S, "params": params, "grips": [], "floor_z": -(half + floor_gap),
            "steps": 24}


def plies_scene(nx=40, ny=32, plies=2, gap=0.75e-3, turn=0.35, perturbation=0.3e-3):
    """A stack of denim plies compiled by the reference's scene compiler (one LayerSpec per ply,
    scene.ts:233-278), each mounted by its collar row (pinned): ply 0 at z = 0, each ply above it
    `gap` higher, 4 nodes narrower and turned by `turn` rad against the one below. The contact
    distance of two denim plies is 0.7 mm (half the summed 0.7 mm thicknesses)."""
    from .material import DENIM
    from .scene import compile_scene, coupon
    specs, sizes = [], []
    for l in range(plies):
        n_u, n_v = nx - 4 * l, ny - 4 * l
        specs.append(coupon(n_u, n_v, material=DENIM, z=l * gap, angle=l * turn, collarRows=1,
                            seed=11 + l, perturbation=perturbation if l == 0 else 0.0))
        sizes.append((n_u, n_v))
    return compile_scene(specs, sizes)


def centre_grip(S, layer, k, size=5):
    """A size x size patch grip at the centre of a layer's lattice (the ReferenceGrip quadratic
    spring, solver.ts:25-37): equal weights, goal at the patch centroid, offsets from it."""
    L = S.layers[layer]
    i0, j0 = (L.nx - size) // 2, (L.ny - size) // 2
    nodes = [L.node_offset + (j0 + b) * L.nx + (i0 + a) for b in range(size) for a in range(size)]
    pos = S.position[nodes]
    goal = pos.mean(0)
    return {"nodes": nodes, "weights": [1.0 / len(nodes)] * len(nodes), "offsets": pos - goal, "goal": goal.copy(), "k": k}


def config_plies(nx=40, ny=32, plies=2, gap=0.75e-3, press=0.1e-3, newton=3, rel_tol=1e-5):
    """The top ply is pressed onto the ones below by a 5 x 5 grip at its centre whose goal moves
    down `press` per substep (12 mm/s at h = 1/120), with cross-layer contact (the design's
    stacked plies, each mounted by its collar)."""
    S = plies_scene(nx, ny, plies, gap)
    params = make_params(h=1 / 120, newton_iterations=newton, pcg_rel_tol=rel_tol, pcg_max_iterations=100,
                         max_coarsest_nodes=400, coarse_sweeps=2, lag_coarse=True, line_search=True, history=False)
    grip = centre_grip(S, plies - 1, k=2e4)
    return {"name": f"plies{plies}x{nx}x{ny}", "scene": S, "params": params, "grips": [grip],
            "velocity": [np.array([0.0, 0.0, -press])], "steps": 24}

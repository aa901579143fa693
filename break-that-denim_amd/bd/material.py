"""Woven material assets in SI units (m, kg, s, N, J).

Restates the reference's material data model (src/simulation/material.ts:18-231): two yarn
families per weave (0 = u / weft / picks, 1 = v / warp / ends), the denim and backing assets,
bundling of several parallel yarns into one lattice span, and the lumped crossing mass.
A material is a plain dict so it can be serialised and hashed like the reference's JSON.
"""
from __future__ import annotations

import copy

FAMILY_U = 0
FAMILY_V = 1

YARN_KEYS = (
    "linearDensity", "eaLow", "eaHigh", "crimpStrain", "eaCompress", "axialViscosity",
    "bendRigidity", "bendViscosity", "strength", "softeningOpening", "width", "height",
    "crimpAmplitude",
)


def _yarn(**kw):
    return dict(kw)


# material.ts:88-138
DENIM = {
    "id": 0,
    "name": "denim-3x1",
    "pitchU": 0.38e-3,
    "pitchV": 0.58e-3,
    "yarns": [
        _yarn(linearDensity=60e-6, eaLow=6, eaHigh=140, crimpStrain=0.03, eaCompress=3,
              axialViscosity=2e-4, bendRigidity=4e-8, bendViscosity=2e-2, strength=8.5,
              softeningOpening=0.4e-3, width=0.4e-3, height=0.16e-3, crimpAmplitude=0.0),
        _yarn(linearDensity=84e-6, eaLow=10, eaHigh=180, crimpStrain=0.025, eaCompress=5,
              axialViscosity=2e-4, bendRigidity=4.5e-8, bendViscosity=2e-2, strength=11,
              softeningOpening=0.4e-3, width=0.7e-3, height=0.26e-3, crimpAmplitude=0.13e-3),
    ],
    "twillRepeat": 4,
    "twillWarpOver": 3,
    "shearStiffness": 60,
    "crossingFriction": 0.35,
    "crossingNormal0": 0.25,
    "crossingWrap": 0.3,
    "thickness": 0.7e-3,
    "clothFriction": 0.45,
    "airDrag": 1.5,
}

# material.ts:141-187
_BACKING_YARN = dict(linearDensity=55e-6, eaLow=5, eaHigh=90, crimpStrain=0.04, eaCompress=0.3,
                     axialViscosity=2e-4, bendRigidity=6e-9, bendViscosity=2e-2, strength=6,
                     softeningOpening=0.25e-3, width=0.52e-3, height=0.34e-3, crimpAmplitude=0.08e-3)
BACKING = {
    "id": 1,
    "name": "brushed-backing",
    "pitchU": 0.5e-3,
    "pitchV": 0.5e-3,
    "yarns": [dict(_BACKING_YARN), dict(_BACKING_YARN)],
    "twillRepeat": 3,
    "twillWarpOver": 2,
    "shearStiffness": 45,
    "crossingFriction": 0.45,
    "crossingNormal0": 0.2,
    "crossingWrap": 0.3,
    "thickness": 1.0e-3,
    "clothFriction": 0.55,
    "airDrag": 1.5,
}


def yarns_per_span(yarn) -> int:
    return int(yarn.get("yarnsPerSpan", 1))


def bundle_weave(m, bu: int, bv: int | None = None):
    """Mechanical material of a lattice whose spans bundle bu warps x bv picks (material.ts:198-226).

    u-family spans (picks) carry bv yarns, v-family spans (warps) carry bu yarns; extensive yarn
    quantities scale with the count, the diagonal shear spring keeps the in-plane shear modulus.
    """
    if bv is None:
        bv = bu
    for b in (bu, bv):
        if int(b) != b or b < 1:
            raise ValueError(f"bundle factors must be positive integers, got {bu}, {bv}")
    if bu == 1 and bv == 1:
        return m

    def yarn(y, b):
        out = dict(y)
        for k in ("linearDensity", "eaLow", "eaHigh", "eaCompress", "axialViscosity", "bendRigidity", "strength"):
            out[k] = b * y[k]
        out["yarnsPerSpan"] = b
        return out

    def shape(a, c):
        return (a * a + c * c) / (a * c)

    out = dict(m)
    out["name"] = f"{m['name']}/b{bu}x{bv}"
    out["pitchU"] = bu * m["pitchU"]
    out["pitchV"] = bv * m["pitchV"]
    out["yarns"] = [yarn(m["yarns"][0], bv), yarn(m["yarns"][1], bu)]
    out["shearStiffness"] = m["shearStiffness"] * shape(bu * m["pitchU"], bv * m["pitchV"]) / shape(m["pitchU"], m["pitchV"])
    out["crossingNormal0"] = bu * bv * m["crossingNormal0"]
    return out


def crossing_mass(m) -> float:
    """Mass lumped at one crossing node: one u-span of weft plus one v-span of warp (material.ts:229-231)."""
    return m["yarns"][0]["linearDensity"] * m["pitchU"] + m["yarns"][1]["linearDensity"] * m["pitchV"]


def linear_sheet_material(pitch: float, e_warp: float, warp_weft_ratio: float, bend_k: float,
                          areal_density: float, shear_modulus: float, name: str = "linear-sheet"):
    """A single-yarn lattice material with linear axial laws, for sheet-like workloads.

    Axial stiffness per span is the membrane stiffness times the pitch (N), bending rigidity per
    span the plate bending stiffness times the pitch (N m^2), the lumped crossing mass is
    areal_density * pitch^2, and the diagonal shear spring of a square cell equals the membrane
    shear modulus (a diagonal pair under simple shear gamma stores 1/2 ks gamma^2 pitch^2).
    Remaining constants (viscosities, friction, drag, strength) are the denim asset's.
    """
    ea_v = e_warp * pitch
    ea_u = ea_v / warp_weft_ratio
    rho = areal_density * pitch / 2.0
    m = copy.deepcopy(DENIM)
    m["name"] = name
    m["id"] = 100
    m["pitchU"] = pitch
    m["pitchV"] = pitch
    for f, ea in ((0, ea_u), (1, ea_v)):
        y = m["yarns"][f]
        y["linearDensity"] = rho
        y["eaLow"] = ea
        y["eaHigh"] = ea
        y["eaCompress"] = ea
        y["bendRigidity"] = bend_k * pitch
    m["shearStiffness"] = shear_modulus
    return m

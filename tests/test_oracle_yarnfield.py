"""Oracle yarn damage field pinned by the reference's fixtures (tests/unit/yarnField.test.ts)."""
import math

import numpy as np
import pytest

import oracle as O
from bd.material import DENIM
from bd.scene import compile_scene, layer_spec, SEG_BROKEN, SEG_SLIPPED, DAMAGE_CUT, DAMAGE_PULLED
from bd.capi import make_params


def yf_scene():  # yarnField.test.ts:7-25
    spec = layer_spec(DENIM, bundle=[9, 6], sizeU=12 * 9 * DENIM["pitchU"], sizeV=10 * 6 * DENIM["pitchV"],
                      collarRows=1, seed=3, perturbation=0)
    return compile_scene([spec])


def make():
    S = yf_scene()
    return S, O.Oracle(S, make_params())


def test_interpolates_uniform_strain_and_ignores_empty_spans():  # yarnField.test.ts:49-57
    S, o = make()
    L = S.layers[0]
    o.strain[:] = 0.02
    assert o.interpolated_strain(1, 40, 30) == pytest.approx(0.02, abs=5e-13)
    o.strain[:] = np.nan
    o.strain[2 * (3 * L.nx + 4) + 1] = 0.05
    c = o.span_crossing(4, 3, 1, 4, 2) - L.field_offset
    assert o.interpolated_strain(1, c % L.K, c // L.K) == pytest.approx(0.05, abs=5e-13)


def test_intact_fraction_and_rupture_conserve_length():  # yarnField.test.ts:59-80
    S, o = make()
    L = S.layers[0]
    i, j, f = 4, 3, 1
    s = 2 * (j * L.nx + i) + f
    dmg = o.damage
    for q in range(3):
        dmg[2 * o.span_crossing(i, j, f, q, q + 1) + f] = DAMAGE_CUT
    dmg[2 * o.span_crossing(i, j, f, 3, 0) + f] = 127
    assert o.couple_spans() == 0
    assert o.intact[s] == pytest.approx((5 + (1 - 127 / 255)) / 9, abs=5e-10)
    assert o.cut_mask[s] == 0b111
    total = o.tails.sum()
    for q in range(9):
        dmg[2 * o.span_crossing(i, j, f, q, 2) + f] = DAMAGE_CUT
    assert o.couple_spans() == 1
    assert o.seg_flags[s] & SEG_BROKEN == SEG_BROKEN
    assert o.intact[s] == 0
    assert o.tails.sum() - total == pytest.approx(o.rest[s], abs=5e-13)


def test_ruptured_span_yarns_part_from_other_side():  # yarnField.test.ts:82-106
    S, o = make()
    i, j, f = 4, 3, 0
    L = S.layers[0]
    P = L.bu
    dmg = o.damage
    cuts = [[0], [P - 1], [4], [1, 6]]
    for q in range(L.bv):
        for p in (cuts[q] if q < len(cuts) else [2]):
            dmg[2 * o.span_crossing(i, j, f, q, p) + f] = DAMAGE_CUT
    before = dmg.copy()
    assert o.couple_spans() == 1
    assert P - P // 2 == 5
    pulled = lambda q: "".join("1" if dmg[2 * o.span_crossing(i, j, f, q, p) + f] == DAMAGE_PULLED else "0" for p in range(P))
    assert pulled(0) == "011110000"
    assert pulled(1) == "000001110"
    assert pulled(2) == "000000000"
    assert pulled(3) == "001110000"
    assert pulled(4) == "000110000"
    changed = np.nonzero(dmg != before)[0]
    assert np.all(changed % 2 == f)


def test_slipped_yarns_part_once():  # yarnField.test.ts:108-129
    S, o = make()
    L = S.layers[0]
    f = 1
    P = L.bv
    for i, site in ((4, 0), (6, 1)):
        s = 2 * (3 * L.nx + i) + f
        o.seg_flags[s] |= SEG_BROKEN | SEG_SLIPPED
        o.fail_site[s] = site
    dmg = o.damage
    dmg[2 * o.span_crossing(6, 3, f, 0, 1) + f] = DAMAGE_CUT
    assert o.couple_spans() == 0
    assert P == 6
    pulled = lambda i, q: "".join("1" if dmg[2 * o.span_crossing(i, 3, f, q, p) + f] == DAMAGE_PULLED else "0" for p in range(P))
    for q in range(L.bu):
        assert pulled(4, q) == "111000"
        assert pulled(6, q) == "000111"
    assert o.seg_flags[2 * (3 * L.nx + 4) + f] & SEG_SLIPPED == 0


def test_overstretched_yarn_separates_at_weakest_crossing_only():  # yarnField.test.ts:131-156
    S, o = make()
    o.strain[:] = 0.5
    o.update_field()
    i, j = 4, 3
    L = S.layers[0]
    dmg = o.damage
    for f in (0, 1):
        B = L.bu if f == 1 else L.bv
        P = L.bv if f == 1 else L.bu
        for q in range(B):
            cuts = 0
            weakest, low = -1, math.inf
            for p in range(P):
                c = o.span_crossing(i, j, f, q, p)
                if dmg[2 * c + f] == DAMAGE_CUT:
                    cuts += 1
                peak = S.field_strength[2 * c + f] / (S.field_area[2 * c + f] / 255)
                if peak < low:
                    low, weakest = peak, c
            assert cuts == 1
            assert dmg[2 * weakest + f] == DAMAGE_CUT


def test_cut_end_pulls_out_by_shear_lag():  # yarnField.test.ts:158-193
    S, o = make()
    L = S.layers[0]
    c0 = L.field_offset + 30 * L.K + 40
    dmg = o.damage
    dmg[2 * c0] = DAMAGE_CUT

    def pull_all(val_u):
        total = 0
        o.strain[:] = 0
        o.strain[0::2] = val_u
        while True:
            n = o.pull_out()
            if n == 0:
                return total
            total += n

    assert pull_all(0.0) == 0
    pull_all(0.06)
    y = L.material["yarns"][0]
    B = L.bv
    comp = (O.axial_tension(y, 0.06) * 0.06 - O.axial_energy_density(y, 0.06)) / B
    p = L.weave["pitchU"]
    n0 = S.crossing_normal0[0] / (L.bu * L.bv)
    mu = L.material["crossingFriction"]
    delta = comp * p / (mu * n0)
    assert O.recoil_length(y, B, 1, 1e3, 0.06, mu, n0, L.material["crossingWrap"], 0, p) == pytest.approx(delta, abs=5e-13)
    peak = float(S.field_strength[2 * c0])
    capped = O.recoil_length(y, B, 1, peak, 0.6, mu, n0, L.material["crossingWrap"], 0, p)
    assert capped == pytest.approx(O.recoil_length(y, B, 1, peak, 0.2, mu, n0, L.material["crossingWrap"], 0, p), abs=5e-13)
    run = min(12, math.floor(delta / p + 0.5))
    assert run > 0
    for d in (-1, 1):
        for k in range(1, run + 1):
            assert dmg[2 * (c0 + d * k)] == DAMAGE_PULLED
        assert dmg[2 * (c0 + d * (run + 1))] < DAMAGE_PULLED
    assert pull_all(0.0) == 0
    assert dmg[2 * (c0 + 1)] == DAMAGE_PULLED


def test_share_factor_runs_and_cap():  # yarnField.test.ts:195-208
    S, o = make()
    L = S.layers[0]
    sep = np.zeros(2 * S.crossing_count, np.uint8)
    reach = o.share_reach(1)
    for k in range(40, 43):
        sep[2 * (L.field_offset + 30 * L.K + k) + 1] = 1
    assert o.share_factor(sep, 1, 43, 30, reach) == pytest.approx(1 + 0.6 * 3 / 2, abs=5e-13)
    assert o.share_factor(sep, 1, 39, 30 + reach, reach) == pytest.approx(1 + 0.6 * 3 / 2, abs=5e-13)
    assert o.share_factor(sep, 1, 43, 30 + reach + 1, reach) == 1
    assert o.share_factor(sep, 1, 45, 30, reach) == 1
    assert o.share_factor(sep, 1, 43, 30, reach, 0.1) == pytest.approx(1.1, abs=5e-13)
    assert O.lib().or_share_cap(2, 0.35, 0.25, 0.3, 0, 11) == pytest.approx(5 * 0.35 * 0.25 / 11, abs=5e-13)


def test_crack_runs_across_bundle_boundary():  # yarnField.test.ts:210-226
    S, o = make()
    L = S.layers[0]
    dmg = o.damage
    k_edge = 4 * L.bu - 1
    m_cut = 3 * L.bv + L.bv // 2 + 1
    dmg[2 * (L.field_offset + m_cut * L.K + k_edge) + 1] = DAMAGE_CUT
    o.couple_spans()
    o.strain[:] = 0.2
    o.update_field()
    reach = o.share_reach(1)
    at = -1
    for m in range(L.M):
        if dmg[2 * (L.field_offset + m * L.K + k_edge + 1) + 1] == DAMAGE_CUT and abs(m - m_cut) <= reach:
            at = m
    assert at >= 0

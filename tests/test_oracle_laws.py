"""Oracle laws pinned by the reference's unit tests (tests/unit/axial.test.ts, laws.test.ts)
and closed forms."""
import math

import numpy as np
import pytest

import oracle as O
from bd.material import DENIM

P = DENIM["yarns"][1]


def fd(f, x, h):
    return (f(x + h) - f(x - h)) / (2 * h)


def test_tension_and_stiffness_are_derivatives():  # axial.test.ts:8-16
    for eps in (-0.01, 0.001, 0.01, 0.02, 0.03, 0.05, 0.08):
        h = 1e-6
        assert fd(lambda e: O.axial_energy_density(P, e), eps, h) == pytest.approx(O.axial_tension(P, eps), abs=5e-6)
        assert fd(lambda e: O.axial_tension(P, e), eps, h) == pytest.approx(O.axial_stiffness(P, eps), abs=5e-4)


def test_strain_at_tension_inverts():  # axial.test.ts:18-20
    for t in (0.01, 0.1, 0.5, 2, 11):
        assert O.axial_tension(P, O.strain_at_tension(P, t)) == pytest.approx(t, abs=5e-10)


def test_cohesive_softening_dissipation():  # axial.test.ts:22-41
    B = 9
    y = dict(P, eaLow=B * P["eaLow"], eaHigh=B * P["eaHigh"], eaCompress=B * P["eaCompress"])
    peak, es = 3, 0.05
    ep = O.strain_at_tension(y, peak * B)
    steps = 20000
    eps = ep + (np.arange(steps) + 0.5) / steps * es
    work = sum((1 - O.yarn_damage(y, B, 1, peak, es, e)) * O.axial_tension(y, e) * es / B / steps for e in eps)
    assert work == pytest.approx(peak * es / 2, abs=5e-7)
    assert O.yarn_damage(y, B, 1, peak, es, 0.99 * ep) == 0
    assert O.yarn_damage(y, B, 1, peak, es, ep + es) == 1
    assert O.yarn_damage(y, B, 0.3, 0.3 * peak, es, 0.99 * ep) == 0
    assert O.yarn_damage(y, B, 0.3, 0.3 * peak, es, ep + es) == 1


def test_damage_monotone():  # axial.test.ts:43-57
    peak, es = 3, 0.05
    ep = O.strain_at_tension(P, peak)
    prev = 0
    for k in range(51):
        d = O.yarn_damage(P, 1, 1, peak, es, ep + k / 50 * es)
        assert d >= prev
        prev = d


A_FOLD = 4 * 1.2e-3 * 1.2e-3


def test_bend_psi_quadratic_and_tan2():  # laws.test.ts:8-15
    q = 1e-4 * A_FOLD
    assert O.bend_psi(q, A_FOLD)[0] / q == pytest.approx(1, abs=5e-4)
    th = 2.0
    s = math.sin(th / 2) ** 2
    assert O.bend_psi(s * A_FOLD, A_FOLD)[0] == pytest.approx(A_FOLD * math.tan(th / 2) ** 2, abs=5e-13)


def test_bend_psi_derivatives_and_c2_joint():  # laws.test.ts:16-27
    a = A_FOLD
    for s in (0.1, 0.5, 0.8, 0.85, 0.95, 1.3):
        q = s * a
        _, d1, d2 = O.bend_psi(q, a)
        assert fd(lambda x: O.bend_psi(x, a)[0], q, 1e-7 * a) / d1 == pytest.approx(1, abs=5e-6)
        assert fd(lambda x: O.bend_psi(x, a)[1], q, 1e-7 * a) / d2 == pytest.approx(1, abs=5e-5)
    q0 = 0.85 * a
    lo = O.bend_psi(q0 * (1 - 1e-12), a)
    hi = O.bend_psi(q0 * (1 + 1e-12), a)
    for k in range(3):
        assert hi[k] / lo[k] == pytest.approx(1, abs=5e-7)


def test_bend_psi_delta():  # laws.test.ts:28-40
    a = A_FOLD
    q0, dq = 0.3 * a, 1e-9 * a
    assert O.bend_psi_delta(q0, q0 + dq, dq, a) / (O.bend_psi(q0, a)[1] * dq) == pytest.approx(1, abs=5e-7)
    for s0, s1 in ((0.2, 0.5), (0.8, 0.9), (0.9, 0.8), (0.95, 3.0), (40, 41), (41, 40)):
        ref = O.bend_psi(s1 * a, a)[0] - O.bend_psi(s0 * a, a)[0]
        assert O.bend_psi_delta(s0 * a, s1 * a, (s1 - s0) * a, a) / ref == pytest.approx(1, abs=5e-10)
    assert O.bend_psi_delta(50 * a, 50 * a, 0, a) == 0


def test_shear_law_branches():  # laws.test.ts:43-59
    ks, Ld = 60, 1.6e-3
    assert O.shear_law(ks, Ld, Ld * 1.1)[1] == ks
    assert O.shear_law(ks, Ld, Ld * (1 - 0.16 + 0.01))[1] == ks
    assert O.shear_law(ks, Ld, Ld * (1 - 0.16 - 0.08 - 0.05))[1] == pytest.approx(40 * ks, abs=5e-10)
    for r in (1.2, 0.9, 1 - 0.16 - 0.04, 1 - 0.16 - 0.16):
        l = r * Ld
        f, k, _ = O.shear_law(ks, Ld, l)
        assert fd(lambda x: O.shear_law(ks, Ld, x)[2], l, 1e-9) / f == pytest.approx(1, abs=5e-6)
        assert fd(lambda x: O.shear_law(ks, Ld, x)[0], l, 1e-9) / k == pytest.approx(1, abs=5e-6)


def test_shear_jam():  # laws.test.ts:61-72
    assert O.shear_jam(1) == pytest.approx(0.16, abs=5e-13)
    g = math.asin(1 - (1 - O.shear_jam(0.5)) ** 2)
    assert 44 < g * 180 / math.pi < 50
    a = 0.05
    while a < 1:
        assert O.shear_jam(a) > O.shear_jam(a + 0.05)
        a += 0.05
    assert O.cell_area(1, 1, 0.2, 0.4) == 1
    assert O.cell_area(0, 0.4, 0.3, 0.3) == pytest.approx(0.3, abs=5e-13)


def test_bend_geometry():  # laws.test.ts:76-109
    L = 3.4e-3
    g = O.bend_geometry(L, L)
    assert g["wa"] == pytest.approx(1, abs=5e-15)
    assert g["wc"] == pytest.approx(1, abs=5e-15)
    assert g["kPerB"] * L ** 3 == pytest.approx(1, abs=5e-13)
    assert g["fold"] / (4 * L * L) == pytest.approx(1, abs=5e-13)

    def curvature(La, Lc, pa, pb, pc):
        g = O.bend_geometry(La, Lc)
        return [g["wa"] * pa[k] - 2 * pb[k] + g["wc"] * pc[k] for k in range(3)]

    for strain in (0, 0.05):
        s = 1 + strain
        r = curvature(3e-3, 4e-3, [0, 0, 0], [3e-3 * s, 0, 0], [7e-3 * s, 0, 0])
        assert math.hypot(*r) < 1e-18
    B, R, La, Lc = 2.4e-7, 0.02, 2.6e-3, 4.1e-3
    at = lambda s: [R * math.sin(s / R), R * (1 - math.cos(s / R)), 0]
    r = curvature(La, Lc, at(-La), at(0), at(Lc))
    g = O.bend_geometry(La, Lc)
    e = 0.5 * B * g["kPerB"] * sum(x * x for x in r)
    assert e / (0.5 * B * (0.5 * (La + Lc)) / (R * R)) == pytest.approx(1, abs=5e-3)


def test_bend_scale():  # laws.test.ts:110-116
    assert O.bend_scale(1, 1, 0, 0, 9) == pytest.approx(1, abs=5e-13)
    assert O.bend_scale(6 / 9, 0.5, 0b111, 0b111, 9) == pytest.approx(0.5, abs=5e-13)
    assert O.bend_scale(4 / 9, 4 / 9, 0b000011111, 0b111110000, 9) == 0


def test_choose_step_cases():  # newtonMultigrid.ts:597-603 decision table
    m = [1.0] * 4
    assert O.choose_step([-1, -2, -0.5, -0.1], m) == 0.5          # conclusive decrease: best alpha
    assert O.choose_step([1, 1, 1, 1], m) == 0                     # every step conclusively increases
    assert O.choose_step([1e-7, -1e-7, 0, 1], m) == 1              # within rounding: full step
    assert O.choose_step([2e-6, 1e-3, 1e-3, 1e-3], [1.0, 1, 1, 1]) == 1  # first change below LINE_NOISE

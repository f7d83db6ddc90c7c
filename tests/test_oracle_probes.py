"""Pins of the oracle's probe restatements by closed forms (no GPU).

TemperatureClassic (src/temperatureClassic.cpp:57-158), pinned also by the
reference's T goldens in test_oracle_smoke_golden.py:
  * two particles of one cell moving with v and -v: T = m v^2 / 2 / |q|, and
    m c^2 (gamma - 1) / |q| above REL_LIMIT = 5e7 m/s;
  * one particle moving along one axis: T = 0; an empty cell: NaN (0/0);
  * the last row and column are never normalised and stay 0.
ParticlesDensity::calc_density (src/particlesDensity.cpp:40-225):
  * every branch spreads the particle over volume v_0 exactly, so
    sum over nodes of density * V_node = sum of mass_array / m_e, with V_node
    the axis disc pi dz dr^2 / 4 at i = 0 and 2 pi i dr^2 dz above;
  * hand values of the near-axis branch (r <= dr/2), whose weight
    (dr^2/2 - r dr + r^2) has no counterpart in the charge weighting.
"""
import numpy as np
import pytest

EL_MASS, EL_CHARGE, C = 9.1e-31, 1.6e-19, 3e8
PI = 3.1415926535897932


def _sim(orc, nr=9, nz=17, dr=5e-4):
    return orc.Sim(nr, nz, (nr - 1) * dr, (nz - 1) * dr, 0.5 * dr / C)


def test_temperature_two_particles(orc):
    s = _sim(orc)
    dr = s.dr
    v = 1.0e6
    # cell (3, 5): r in (3 dr, 4 dr], z in (5 dz, 6 dz]
    r = np.array([3.4 * dr, 3.7 * dr, 6.5 * dr])
    z = np.array([5.2 * dr, 5.9 * dr, 2.5 * dr])
    s.add_plasma(-1, 1.0, r, z, np.array([v, -v, 0.0]), np.zeros(3), np.array([0.0, 0.0, 2.0 ** 20]), np.ones(3))
    t = s.temperature(0)
    assert t[3, 5] == pytest.approx(EL_MASS * v * v / 2 / EL_CHARGE, rel=1e-15)
    assert t[6, 2] == 0.0               # one particle along z: S^2 - |V|^2 = 0 exactly
    assert np.isnan(t[0, 0])            # empty cell
    assert np.all(t[-1, :] == 0) and np.all(t[:, -1] == 0)


def test_temperature_relativistic_branch(orc):
    s = _sim(orc)
    dr = s.dr
    v = 1.5e8
    r = np.array([2.5 * dr, 2.6 * dr])
    z = np.array([4.5 * dr, 4.6 * dr])
    s.add_plasma(1, 1836.0, r, z, np.zeros(2), np.array([v, -v]), np.zeros(2), np.ones(2))
    t = s.temperature(0)
    m = 1836.0 * EL_MASS
    gamma = 1.0 / np.sqrt(1.0 - v * v / C ** 2)
    assert t[2, 4] == pytest.approx(m * C ** 2 * (gamma - 1) / EL_CHARGE, rel=1e-14)
    # the classical formula would be 13% lower: the branch is taken
    assert t[2, 4] > 1.1 * m * v * v / 2 / EL_CHARGE


def _node_volumes(nr, nz, dr, dz):
    i = np.arange(nr, dtype=np.float64)[:, None]
    v = PI * dz * dr * dr * 2.0 * i * np.ones((1, nz))
    v[0, :] = PI * dz * dr * dr / 4.0
    return v


@pytest.mark.parametrize("mass", [1.0, 1836.0])
def test_density_conserves_particles(orc, mass):
    rng = np.random.default_rng(5)
    s = _sim(orc, nr=12, nz=20)
    dr, dz = s.dr, s.dz
    n = 4000
    # every branch: r <= dr/2, dr/2 < r <= dr, r > dr (interior cells only in z)
    r = np.concatenate([rng.uniform(0.01, 0.5, 500), rng.uniform(0.5, 1.0, 500), rng.uniform(1.0, 10.9, n - 1000)]) * dr
    z = rng.uniform(0.6, 18.4, n) * dz
    w = rng.uniform(1e3, 1e5, n)
    s.add_plasma(-1, mass, r, z, np.zeros(n), np.zeros(n), np.zeros(n), w)
    d = s.density(0)
    total = np.sum(d * _node_volumes(s.nr, s.nz, dr, dz))
    assert total == pytest.approx(np.sum(mass * w), rel=1e-12)
    # and per branch
    for lo, hi in ((0, 500), (500, 1000), (1000, n)):
        s2 = _sim(orc, nr=12, nz=20)
        s2.add_plasma(-1, mass, r[lo:hi], z[lo:hi], np.zeros(hi - lo), np.zeros(hi - lo), np.zeros(hi - lo), w[lo:hi])
        t2 = np.sum(s2.density(0) * _node_volumes(s2.nr, s2.nz, dr, dz))
        assert t2 == pytest.approx(np.sum(mass * w[lo:hi]), rel=1e-12), (lo, hi)


def test_density_near_axis_hand_values(orc):
    s = _sim(orc)
    dr, dz = s.dr, s.dz
    # r = dr/4 in cell (0, 3), z at 3.5 dz: dz1 = dz2 = dz/2
    s.add_plasma(-1, 1.0, np.array([0.25 * dr]), np.array([3.5 * dz]), np.zeros(1), np.zeros(1), np.zeros(1),
                 np.array([1.0]))
    d = s.density(0)
    r = 0.25 * dr
    v0 = PI * dz * (2 * r * r + dr * dr / 2)
    ro_p = 1.0 / v0
    # axis node: pi dz1 (dr^2/2 - r dr + r^2) / (pi dz dr^2 / 4) = (5/4) dz1 / dz
    assert d[0, 3] == pytest.approx(ro_p * 1.25 * 0.5, rel=1e-14)
    assert d[0, 4] == pytest.approx(ro_p * 1.25 * 0.5, rel=1e-14)
    # ring node 1: pi dz1 ((r + dr/2)^2 - dr^2/4) / (2 pi dr^2 dz) = dz1 (r^2 + r dr) / (2 dr^2 dz)
    w1 = 0.5 * (r * r + r * dr) / (2 * dr * dr)
    assert d[1, 3] == pytest.approx(ro_p * w1, rel=1e-14)
    assert d[1, 4] == pytest.approx(ro_p * w1, rel=1e-14)
    assert np.count_nonzero(d) == 4

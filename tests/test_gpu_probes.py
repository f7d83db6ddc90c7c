"""Device probes (pdp3_temperature, pdp3_density) against the oracle's
TemperatureClassic (src/temperatureClassic.cpp:57-158) and
ParticlesDensity::calc_density (src/particlesDensity.cpp:40-225) on the same
particle state.

* FP64 plain form: the same empty cells (NaN) and last row / column (0), and
  the values within 1e-12 normwise (the per-cell sums use FP64 atomics, so
  they differ from the serial sums by round-off of the order; one-particle
  cells hold S^2 - |V|^2 of a few ulps);
* FP32 form: the probes read the cell form in place, so they see the FP32
  rounding of offsets and momenta: 1e-5 normwise for T, 1e-6 for the density.
"""
import numpy as np
import pytest

from helpers import normwise

pytestmark = pytest.mark.gpu


def _state(seed, nr, nz, n):
    rng = np.random.default_rng(seed)
    dr = 5e-4
    r = dr / 2 + rng.uniform(0, 1, n) * (nr - 2) * dr
    r[: n // 50] = rng.uniform(0.05, 1.0, n // 50) * dr  # the axis cell
    # the plasma fills the first 3/4 in z, a thin sprinkling the rest: empty
    # and one-particle cells
    z = dr / 2 + rng.uniform(0, 0.75, n) * (nz - 2) * dr
    z[-n // 200:] = dr / 2 + rng.uniform(0.75, 1.0, n // 200) * (nz - 2) * dr
    vth = np.sqrt(1.0 * 1.6e-19 / 9.1e-31)
    v = rng.normal(0, vth, (3, n))
    v[:, : n // 20] *= 300.0  # some particles above the 5e7 m/s classical limit
    v = np.clip(v, -1.5e8, 1.5e8)
    w = rng.uniform(1e4, 1e6, n)
    return r, z, v, w


@pytest.mark.parametrize("precision", [0, 1])
def test_probes_match_oracle(orc, precision):
    from pdp3_amd import capi

    nr, nz, n = 33, 129, 60000
    dr = 5e-4
    dt = 0.5 * dr / 3e8
    r, z, v, w = _state(3, nr, nz, n)
    print("probes: 33x129, 6e4 particles, precision", precision, flush=True)
    gpu = capi.Simulation(nr, nz, (nr - 1) * dr, (nz - 1) * dr, dt, precision=precision)
    gpu.add_plasma(-1, 1.0, r, z, v[0], v[1], v[2], w)
    gpu.add_plasma(1, 1836.0, r[::3], z[::3], v[0, ::3] / 40, v[1, ::3] / 40, v[2, ::3] / 40, w[::3])
    gpu.step(7)
    ora = orc.Sim(nr, nz, (nr - 1) * dr, (nz - 1) * dr, dt)
    for j, (q, m) in enumerate(((-1, 1.0), (1, 1836.0))):
        st = gpu.download(j)
        ora.add_plasma(q, m, st["r"], st["z"], st["vr"], st["vphi"], st["vz"], st["w"], st["alive"])
    t_bar, d_bar = (1e-12, 1e-13) if precision == 0 else (1e-5, 1e-6)
    for j in range(2):
        t, t_ref = gpu.temperature(j), ora.temperature(j)
        assert np.array_equal(np.isnan(t), np.isnan(t_ref)), j
        assert np.isnan(t_ref).any() and (np.nan_to_num(t_ref) > 0).sum() > 1000
        assert np.all(t[-1, :] == 0) and np.all(t[:, -1] == 0)
        fin = ~np.isnan(t_ref)
        e_t = normwise(t[fin], t_ref[fin])
        d, d_ref = gpu.density(j), ora.density(j)
        e_d = normwise(d, d_ref)
        print("species", j, "T", e_t, "density", e_d)
        assert e_t <= t_bar, (j, e_t)
        assert e_d <= d_bar, (j, e_d)
    gpu.close()


def test_probe_arguments(orc):
    from pdp3_amd import capi

    dr = 5e-4
    gpu = capi.Simulation(9, 17, 8 * dr, 16 * dr, 1e-12)
    with pytest.raises(capi.Pdp3Error):
        gpu.temperature(0)  # no such species
    gpu.add_plasma(-1, 1.0, [2.5 * dr], [3.5 * dr], [1e5], [0.0], [0.0], [1.0])
    t = gpu.temperature(0)
    assert t[2, 3] == 0.0 and np.isnan(t[0, 0])
    d = gpu.density(0)
    assert d.shape == (9, 17) and np.count_nonzero(d) == 4
    gpu.close()

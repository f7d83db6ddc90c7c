"""Charge conservation of the oracle: the Gauss-law residual div E - rho/eps0
keeps its initial value to round-off while plasma particles cross cells,
reflect at the walls and the axis, with either pusher, and while a bunch
streams through the grid (outside its injection nodes).

This pins the current deposit of the oracle (src/particles.cpp:333-988)
against its charge weighting (src/particles.cpp:101-226) and the E update
(src/eField.cpp:85-170): a dropped or mis-signed term in any segment integral
or a wrong face in the cell splitting breaks the balance at the nodes it
touches.  The reference itself never evaluates the residual; without the PML
layer it stays at 1e-14 of max |rho|/eps0 over 100 steps.
"""
import numpy as np
import pytest

from helpers import gauss_residual_np, oracle_rho_all


def _setup(orc, pusher, temperature_ev, bunch, seed=2):
    rng = np.random.default_rng(seed)
    nr, nz = 33, 129
    dr = dz = 5e-4
    dt = 0.5 * dr / 3e8
    s = orc.Sim(nr, nz, (nr - 1) * dr, (nz - 1) * dz, dt, pusher=pusher)
    n = 20000
    r = dr / 2 + rng.uniform(0.0, 1.0, n) * ((nr - 2) * dr)
    z = dz / 2 + rng.uniform(0.0, 1.0, n) * ((nz - 2) * dz)
    vth = np.sqrt(temperature_ev * 1.6e-19 / 9.1e-31)
    v = np.clip(rng.normal(0, vth, (3, n)), -1.5e8, 1.5e8)
    s.add_plasma(-1, 1.0, r, z, v[0], v[1], v[2], np.full(n, 1e7))
    if bunch:
        nb = 5000
        rb = dr / 2 + rng.uniform(0.0, 1.0, nb) * 4 * dr
        s.add_bunch(-1, 1.0, rb, np.full(nb, 1e7), rng.uniform(0, 1, nb), 20 * dt, 0.0, 2.7e8, 0)
    return s


def _residual(orc, s):
    return gauss_residual_np(s.field("E_r"), s.field("E_z"), oracle_rho_all(orc, s), s.dr, s.dz)


@pytest.mark.parametrize("pusher,temp,bunch", [(0, 1e3, False), (1, 1e4, False), (0, 1e3, True)])
def test_oracle_conserves_gauss_residual(orc, pusher, temp, bunch):
    s = _setup(orc, pusher, temp, bunch)
    R0 = _residual(orc, s)
    scale = np.max(np.abs(oracle_rho_all(orc, s))) / 8.85e-12
    moved = s.get_species(s.species[0])["z"].copy()
    s.step(100)
    dR = np.abs(_residual(orc, s) - R0) / scale
    kmin = 3 if bunch else 1   # a bunch creates charge at its injection nodes k = 1, 2
    assert dR[:, kmin:].max() < 1e-12, dR[:, kmin:].max()
    # the check is not vacuous: particles did cross cells and the field is not zero
    z1 = s.get_species(s.species[0])["z"]
    assert np.mean(np.floor(moved / s.dz) != np.floor(z1 / s.dz)) > 0.05
    assert np.max(np.abs(s.field("E_z"))) > 0
    if bunch:
        assert dR[:, 1:3].max() > 1e-3   # injection does create charge there


def test_pml_layer_breaks_the_balance(orc):
    # the conductivity layer damps E without a current: the residual moves there
    s = _setup(orc, 0, 1e3, False)
    s.set_pml(0.1, 0.1, 0.1, 1e-5, 7e-2)
    R0 = _residual(orc, s)
    scale = np.max(np.abs(oracle_rho_all(orc, s))) / 8.85e-12
    s.step(100)
    dR = np.abs(_residual(orc, s) - R0) / scale
    assert dR.max() > 1e-6

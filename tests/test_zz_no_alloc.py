"""pdp3_gauss_residual_noalloc (the library's fixed device array instead of
allocated buffers) against pdp3_gauss_residual, called through the C ABI.

Both run the same kernels; the charge weighting sums with FP64 atomics, so
rho and the residual agree to FP64 round-off of the summation order: the
bar is 1e-13 normwise.  Two calls in a row, the second on a smaller grid (the
array keeps the first call's values beyond the second grid), and a grid above
PDP3_NOALLOC_MAX_NODES is refused with PDP3_ERR_BAD_ARG.
"""
import numpy as np
import pytest

from helpers import normwise

pytestmark = pytest.mark.gpu


def _sim(capi, nr, nz, precision, seed):
    dr = 5e-4
    dt = 0.5 * dr / 3e8
    s = capi.Simulation(nr, nz, (nr - 1) * dr, (nz - 1) * dr, dt, precision=precision)
    rng = np.random.default_rng(seed)
    n = 20 * (nr - 1) * (nz - 1)
    r = dr / 2 + rng.uniform(0.0, 1.0, n) * ((nr - 2) * dr)
    z = dr / 2 + rng.uniform(0.0, 1.0, n) * ((nz - 2) * dr)
    v = np.clip(rng.normal(0, 1e6, (3, n)), -1.5e8, 1.5e8)
    s.add_plasma(-1, 1.0, r, z, v[0], v[1], v[2], np.full(n, 1e7))
    nb = 2000
    s.add_bunch(-1, 1.0, dr / 2 + rng.uniform(0, 1, nb) * 4 * dr, np.full(nb, 1e7), rng.uniform(0, 1, nb),
                10 * dt, 0.0, 2.7e8, 0)
    return s


def _compare(sim):
    R0, rho0 = sim.gauss_residual()
    R1, rho1 = sim.gauss_residual_noalloc()
    assert np.max(np.abs(rho0)) > 0
    assert normwise(rho1, rho0) <= 1e-13, normwise(rho1, rho0)
    assert normwise(R1, R0) <= 1e-13, normwise(R1, R0)


def test_noalloc_matches_allocating_entry_point():
    from pdp3_amd import capi

    print("noalloc: 65x257 FP64 then 33x129 FP32", flush=True)
    big = _sim(capi, 65, 257, capi.PREC_FP64, 1)
    big.step(12)
    _compare(big)
    small = _sim(capi, 33, 129, capi.PREC_FP32, 2)
    small.step(12)
    _compare(small)
    # the first context again, after the smaller grid used the array
    big.step(3)
    _compare(big)
    big.close()
    small.close()


def test_noalloc_rejects_grids_above_the_limit():
    from pdp3_amd import capi

    nr, nz = 1025, 4609
    assert nr * nz > capi.NOALLOC_MAX_NODES
    dr = 5e-4
    s = capi.Simulation(nr, nz, (nr - 1) * dr, (nz - 1) * dr, 0.5 * dr / 3e8)
    R = np.empty((nr, nz))
    st = s.L.pdp3_gauss_residual_noalloc(s.h, R.ctypes.data_as(capi._D), None)
    assert st == capi.ERR_BAD_ARG
    assert b"PDP3_NOALLOC_MAX_NODES" in s.L.pdp3_last_error(s.h)
    s.close()

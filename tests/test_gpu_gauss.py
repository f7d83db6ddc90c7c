"""Charge conservation on the GPU: pdp3_gauss_residual against the oracle and
the drift of the residual over 100 steps.

* the device charge density of all species equals the oracle's charge
  weighting (normwise 1e-14: FP64 atomics sum in another order), and the
  device residual equals the flux-form residual of the device's own E and rho;
* FP64 plain form: the residual keeps its initial value to round-off (<=1e-12
  of max |rho|/eps0) with either pusher, with a bunch outside its injection
  nodes, exactly as the oracle does;
* FP32 form: the deposit conserves charge to FP32 round-off of the particle
  arithmetic; the drift bound is stated here.
"""
import numpy as np
import pytest

from helpers import gauss_residual_np, normwise, oracle_rho_all

pytestmark = pytest.mark.gpu

EPS0 = 8.85e-12


def _state(seed, nr, nz, temperature_ev, n):
    rng = np.random.default_rng(seed)
    dr = dz = 5e-4
    r = dr / 2 + rng.uniform(0.0, 1.0, n) * ((nr - 2) * dr)
    z = dz / 2 + rng.uniform(0.0, 1.0, n) * ((nz - 2) * dz)
    vth = np.sqrt(temperature_ev * 1.6e-19 / 9.1e-31)
    v = np.clip(rng.normal(0, vth, (3, n)), -1.5e8, 1.5e8)
    nb = 5000
    rb = dr / 2 + rng.uniform(0.0, 1.0, nb) * 4 * dr
    return dict(r=r, z=z, v=v, w=np.full(n, 1e7), rb=rb, wb=np.full(nb, 1e7), ub=rng.uniform(0, 1, nb))


def _build(mod, st, nr, nz, pusher, bunch, **kw):
    dr = 5e-4
    dt = 0.5 * dr / 3e8
    s = mod(nr, nz, (nr - 1) * dr, (nz - 1) * dr, dt, pusher=pusher, **kw)
    s.add_plasma(-1, 1.0, st["r"], st["z"], st["v"][0], st["v"][1], st["v"][2], st["w"])
    if bunch:
        s.add_bunch(-1, 1.0, st["rb"], st["wb"], st["ub"], 20 * dt, 0.0, 2.7e8, 0)
    return s


def test_rho_and_residual_match(orc):
    from pdp3_amd import capi

    nr, nz = 65, 257
    st = _state(5, nr, nz, 1e3, 100000)
    print("gauss: rho and residual vs oracle, 65x257, 1e5 particles", flush=True)
    gpu = _build(capi.Simulation, st, nr, nz, 0, True)
    ora = _build(orc.Sim, st, nr, nz, 0, True)
    for steps in (0, 30):
        gpu.step(steps) if steps else None
        ora.step(steps) if steps else None
        R, rho = gpu.gauss_residual()
        rho_o = oracle_rho_all(orc, ora)
        assert normwise(rho, rho_o) <= 1e-14, normwise(rho, rho_o)
        R_np = gauss_residual_np(gpu.field("E_r"), gpu.field("E_z"), rho, gpu.dr, gpu.dz)
        assert normwise(R, R_np) <= 1e-14, normwise(R, R_np)


@pytest.mark.parametrize("pusher,bunch", [(0, False), (1, False), (0, True)])
def test_fp64_residual_drift(pusher, bunch):
    from pdp3_amd import capi

    nr, nz = 65, 257
    st = _state(7, nr, nz, 1e3, 160000)
    print(f"gauss drift fp64 pusher={pusher} bunch={bunch}", flush=True)
    gpu = _build(capi.Simulation, st, nr, nz, pusher, bunch)
    R0, rho0 = gpu.gauss_residual()
    scale = np.max(np.abs(rho0)) / EPS0
    gpu.step(100)
    R1, _ = gpu.gauss_residual()
    kmin = 3 if bunch else 1
    drift = np.max(np.abs(R1 - R0)[:, kmin:]) / scale
    print("drift", drift)
    assert drift <= 1e-12
    assert np.max(np.abs(gpu.field("E_z"))) > 0


def test_fp32_residual_drift():
    from pdp3_amd import capi

    nr, nz = 65, 257
    st = _state(7, nr, nz, 1e3, 160000)
    print("gauss drift fp32", flush=True)
    gpu = _build(capi.Simulation, st, nr, nz, 0, False, precision=capi.PREC_FP32)
    R0, rho0 = gpu.gauss_residual()
    scale = np.max(np.abs(rho0)) / EPS0
    gpu.step(100)
    R1, _ = gpu.gauss_residual()
    drift = np.max(np.abs(R1 - R0)) / scale
    print("fp32 drift", drift)
    # FP32 offsets carry 24 bits per coordinate: the charge moved through a
    # face and the change of the two node weights agree to ~1e-7 per crossing
    assert drift <= 1e-5

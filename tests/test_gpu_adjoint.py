"""Adjoint gradients against central finite differences of the FP64 reference.

A reduced copy of the cfg5 problem (32 x 256 cells, 200 steps, checkpoints
every 30 steps so that the last segment is partial, 16 bump amplitudes and
the ramp slope).  The gradient of the axis E_z^2 objective comes from the
library's adjoint, chained through torch (PicObjective).  The reference is
the central difference (relative step 1e-4) of the same objective computed
by the oracle, the deterministic CPU restatement of the reference loop: the
GPU's FP64 form sums its deposits with atomics in a varying order, and the
reference's J_r formula amplifies that round-off (to ~1e-9 relative per 100
steps with a beam), which a 1e-4 step would turn into ~1e-4 noise.
FP64 forward: 1e-4 relative on every component above 1% of the largest;
FP32 forward (the production form): 2e-2 relative, its forward error.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem():
    from pdp3_amd import adjoint as adj

    spec = adj.cfg5_spec(cells_r=32, cells_z=256, steps=200, checkpoint_every=30, window_steps=60, ppc=4,
                         bunch_macro=4000, window_frac=(0.2, 0.4))
    return adj.BumpRampProblem(spec)


def oracle_objective(orc, prob, theta):
    """phi(theta) by the oracle, from the same initial state"""
    s = prob.spec
    wp, wb = (w.detach().numpy() for w in prob.weights(theta))
    sim = orc.Sim(prob.n_grid_r, prob.n_grid_z, s.cells_r * s.dr, s.cells_z * s.dr, prob.dt, threads=8)
    n = s.pml_cells
    sim.set_pml((n + 0.5) / prob.n_grid_z, (n + 0.5) / prob.n_grid_z, (n + 0.5) / prob.n_grid_r, 1e-5, 7e-2)
    st = prob.plasma
    sim.add_plasma(-1, 1.0, st["r"], st["z"], st["vr"], st["vphi"], st["vz"], wp, st["alive"])
    nb = len(prob.bunch_r)
    zero = np.zeros(nb)
    sim.add_bunch_loaded(-1, 1.0, prob.bunch_r, prob.bunch_z, zero, zero, np.full(nb, s.velocity), wb)
    z_lo, z_hi = s.window
    dz = sim.dz
    k0 = max(int(np.ceil(z_lo / dz - 0.5)), 0)
    k1 = min(int(np.floor(z_hi / dz - 0.5)), prob.n_grid_z - 2)
    phi = 0.0
    for step in range(s.steps):
        sim.step(1)
        if step >= s.steps - s.window_steps:
            ez = sim.field("E_z")[0, k0:k1 + 1]
            phi += prob.dt * dz * float(np.sum(ez * ez))
    return phi


@pytest.fixture(scope="module")
def problem_and_fd(orc):
    prob = _problem()
    th = prob.theta0()
    print("adjoint: central differences of the oracle, 17 parameters", flush=True)
    fd = np.zeros(len(th))
    for j in range(len(th)):
        h = 1e-4 * abs(float(th[j]))
        tp, tm = th.clone(), th.clone()
        tp[j] += h
        tm[j] -= h
        fd[j] = (oracle_objective(orc, prob, tp) - oracle_objective(orc, prob, tm)) / (2 * h)
    phi0 = oracle_objective(orc, prob, th)
    return prob, th, fd, phi0


def _check(g, fd, tol):
    big = np.abs(fd) > 1e-2 * np.max(np.abs(fd))
    rel = np.abs(g[big] - fd[big]) / np.abs(fd[big])
    print("fd     ", np.array2string(fd, precision=9))
    print("adjoint", np.array2string(g, precision=9))
    print("max rel", rel.max(), "over", big.sum(), "components")
    assert big.sum() >= 5
    assert rel.max() <= tol, rel


def test_adjoint_fp64_matches_fd(problem_and_fd):
    from pdp3_amd import adjoint as adj
    from pdp3_amd import capi

    prob, th, fd, phi0 = problem_and_fd
    theta = th.clone().requires_grad_(True)
    phi = adj.PicObjective.apply(theta, prob, capi.PREC_FP64)
    phi.backward()
    print("phi gpu", float(phi), "oracle", phi0)
    assert abs(float(phi) - phi0) <= 1e-6 * abs(phi0)
    _check(theta.grad.numpy(), fd, 1e-4)


def test_adjoint_fp32_matches_fd(problem_and_fd):
    from pdp3_amd import adjoint as adj
    from pdp3_amd import capi

    prob, th, fd, phi0 = problem_and_fd
    theta = th.clone().requires_grad_(True)
    phi = adj.PicObjective.apply(theta, prob, capi.PREC_FP32)
    phi.backward()
    print("phi fp32", float(phi), "oracle", phi0)
    _check(theta.grad.numpy(), fd, 2e-2)

"""Adjoint gradients against central finite differences of the FP64 reference.

The reference is the central difference (relative step 1e-4) of the axis
E_z^2 objective computed by the oracle, the deterministic CPU restatement of
the reference loop: the GPU's FP64 form sums its deposits with atomics in a
varying order, and the reference's J_r formula amplifies that round-off.  The
gradient comes from the library's adjoint, chained through torch
(PicObjective), on reduced copies of the cfg5 problem (32 x 256 cells, 200
steps, checkpoints every 30 steps so that the last segment is partial, 16 bump
amplitudes and the ramp slope).

The reference loop is only piecewise smooth in the parameters, and its
deposit formulas lose digits for tiny displacements:
  * a particle reflected at the axis or a wall one step earlier receives its E
    kick after instead of before the velocity flip (a jump of 2 q E dt / m);
  * paths with |dr| << r (a relativistic bunch on a purely axial path, a cold
    plasma starting from rest) evaluate J as differences of O(r dz) terms, and
    the MNZL flushes switch branches.
Difference quotients see these (they move with the step: for a cold plasma
by up to 8% between steps of 1e-2 and 1e-4 of a parameter, while the adjoint
stays put); the adjoint, a derivative along the trajectory, does not.  Hence:
  * the smooth problem (thermal plasma kept three cells away from the axis and
    the walls, a bunch outside the innermost cell with a small radial velocity
    spread): 1e-4 normwise and on the components above a quarter of the
    largest, FP64 forward; 2e-2 for the FP32 forward (its forward error);
  * the realistic problem (plasma down to the axis, a purely axial bunch):
    1e-3 normwise (1e-2 per dominant component), FP64 forward, and 2e-2 for
    the FP32 forward.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(smooth=False):
    from pdp3_amd import adjoint as adj

    extra = dict(plasma_margin_cells=3, bunch_r_min=5e-4, bunch_vr_spread=1e4) if smooth else {}
    spec = adj.cfg5_spec(cells_r=32, cells_z=256, steps=200, checkpoint_every=30, window_steps=60, ppc=4,
                         bunch_macro=4000, window_frac=(0.2, 0.4), **extra)
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
    sim.add_bunch_loaded(-1, 1.0, prob.bunch_r, prob.bunch_z, prob.bunch_vr, zero, np.full(nb, s.velocity), wb)
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


def _fd(orc, prob):
    th = prob.theta0()
    fd = np.zeros(len(th))
    for j in range(len(th)):
        h = 1e-4 * abs(float(th[j]))
        tp, tm = th.clone(), th.clone()
        tp[j] += h
        tm[j] -= h
        fd[j] = (oracle_objective(orc, prob, tp) - oracle_objective(orc, prob, tm)) / (2 * h)
    return prob, th, fd, oracle_objective(orc, prob, th)


@pytest.fixture(scope="module")
def problem_and_fd(orc):
    print("adjoint: central differences of the oracle, 17 parameters", flush=True)
    return _fd(orc, _problem())


@pytest.fixture(scope="module")
def smooth_problem_and_fd(orc):
    print("adjoint (smooth problem): central differences of the oracle, 17 parameters", flush=True)
    return _fd(orc, _problem(smooth=True))


def test_adjoint_fp64_smooth_problem(smooth_problem_and_fd):
    from pdp3_amd import adjoint as adj
    from pdp3_amd import capi

    prob, th, fd, phi0 = smooth_problem_and_fd
    theta = th.clone().requires_grad_(True)
    phi = adj.PicObjective.apply(theta, prob, capi.PREC_FP64)
    phi.backward()
    print("phi gpu", float(phi), "oracle", phi0)
    assert abs(float(phi) - phi0) <= 1e-6 * abs(phi0)
    _check(theta.grad.numpy(), fd, 1e-4, 1e-4)


def test_adjoint_fp32_smooth_problem(smooth_problem_and_fd):
    from pdp3_amd import adjoint as adj
    from pdp3_amd import capi

    prob, th, fd, phi0 = smooth_problem_and_fd
    theta = th.clone().requires_grad_(True)
    phi = adj.PicObjective.apply(theta, prob, capi.PREC_FP32)
    phi.backward()
    print("phi fp32", float(phi), "oracle", phi0)
    _check(theta.grad.numpy(), fd, 2e-2, 2e-2)


def _check(g, fd, tol_norm, tol_dominant):
    scale = np.max(np.abs(fd))
    norm = np.max(np.abs(g - fd)) / scale
    rel = np.abs(g - fd) / np.abs(fd)
    dom = np.abs(fd) >= 0.25 * scale
    print("fd     ", np.array2string(fd, precision=9))
    print("adjoint", np.array2string(g, precision=9))
    print("relative per component", np.array2string(rel, precision=2))
    print("normwise", norm, "dominant max rel", rel[dom].max(), "over", dom.sum())
    assert np.all(np.sign(g[np.abs(fd) > 0.05 * scale]) == np.sign(fd[np.abs(fd) > 0.05 * scale]))
    assert norm <= tol_norm, norm
    assert rel[dom].max() <= tol_dominant, rel[dom]


def test_adjoint_fp64_matches_fd(problem_and_fd):
    from pdp3_amd import adjoint as adj
    from pdp3_amd import capi

    prob, th, fd, phi0 = problem_and_fd
    theta = th.clone().requires_grad_(True)
    phi = adj.PicObjective.apply(theta, prob, capi.PREC_FP64)
    phi.backward()
    print("phi gpu", float(phi), "oracle", phi0)
    assert abs(float(phi) - phi0) <= 1e-6 * abs(phi0)
    _check(theta.grad.numpy(), fd, 1e-3, 1e-2)


def test_adjoint_fp32_matches_fd(problem_and_fd):
    from pdp3_amd import adjoint as adj
    from pdp3_amd import capi

    prob, th, fd, phi0 = problem_and_fd
    theta = th.clone().requires_grad_(True)
    phi = adj.PicObjective.apply(theta, prob, capi.PREC_FP32)
    phi.backward()
    print("phi fp32", float(phi), "oracle", phi0)
    _check(theta.grad.numpy(), fd, 2e-2, 2e-2)


def test_batched_gradients_match_single_runs():
    """pdp3_adjoint_run_batch (members concurrent on their own streams) gives
    each member the gradient of its own single run (FP64 form: the runs agree
    to round-off over the 200 steps of the reduced problem)."""
    import torch

    from pdp3_amd import adjoint as adj
    from pdp3_amd import capi

    prob = _problem()
    th = prob.theta0()
    rng = np.random.default_rng(3)
    thetas = [th * torch.tensor(rng.uniform(0.8, 1.2, len(th))) for _ in range(3)]
    print("batched gradient of 3 members vs single runs", flush=True)
    phis, grads = prob.batch_gradients(thetas, capi.PREC_FP64)
    for m, t in enumerate(thetas):
        theta = t.clone().requires_grad_(True)
        phi = adj.PicObjective.apply(theta, prob, capi.PREC_FP64)
        phi.backward()
        g = theta.grad.numpy()
        e_phi = abs(phis[m] - float(phi)) / abs(float(phi))
        e_g = np.max(np.abs(grads[m] - g)) / np.max(np.abs(g))
        print("member", m, "phi rel diff", e_phi, "gradient normwise diff", e_g)
        assert e_phi <= 1e-10 and e_g <= 1e-8, (m, e_phi, e_g)

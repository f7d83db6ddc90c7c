"""The adjoint proved without finite differences: the dot-product test.

pdp3_tangent_run carries weight tangents forward through the same loop
(forward-mode dual numbers for the particle step, the context's own linear
field updates for the tangent grids), so it computes J v for the Jacobian J of
phi with respect to the particle weights; the adjoint computes J^T.  For
random directions v:

    <dphi/dw (adjoint), v>  ==  dphi along v (tangent-linear)

to round-off of the sums, provided both linearise around one trajectory.
Directions 0-1 are random directions in the 17 parameters (bump amplitudes
and ramp slope; w(theta) is linear), directions 2-3 random per-particle
weight changes.

* reduced problem (32 x 256 cells, 200 steps), separate adjoint and tangent
  runs, both forms (the FP32 form's adjoint and tangent are both the FP64
  step's linearisation around the FP32 trajectory): the two runs' trajectories
  agree to round-off over 200 steps, and so does the identity (1e-10 / 1e-8);
* full cfg5 (128 x 1024 cells, 2.7e6 particles, 1500 steps), FP64: two runs
  drift apart (FP64 atomics sum in a varying order and the loop amplifies
  it, phi moves by ~1e-8 relative), so the tangents ride along the adjoint's
  own forward sweep (pdp3_adjoint_tangent_run); with a checkpoint every step
  the adjoint linearises around exactly those states and the identity holds
  to 1e-8; with checkpoints every 50 steps the recomputed segments deviate
  by amplified round-off, and the identity holds to 1e-4.

Also here: central differences of the oracle converge to the adjoint as the
step shrinks, on the smooth reduced problem.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _directions(prob, theta, rng):
    """4 weight-tangent directions: (per species) arrays of shape (4, n)"""
    import torch

    n_p = len(prob.plasma["r"])
    n_b = len(prob.bunch_r)
    dp = np.zeros((4, n_p))
    db = np.zeros((4, n_b))
    for d in range(2):
        dth = torch.tensor(rng.standard_normal(len(theta)), dtype=torch.float64)
        _, (tp, tb) = torch.func.jvp(lambda t: tuple(prob.weights(t)), (theta,), (dth,))
        dp[d], db[d] = tp.numpy(), tb.numpy()
    w_p, w_b = (w.detach().numpy() for w in prob.weights(theta))
    for d in range(2, 4):
        dp[d] = w_p * rng.standard_normal(n_p)
        db[d] = w_b * rng.standard_normal(n_b)
    return [dp, db]


def _dot_product(prob, precision, seed):
    from pdp3_amd import capi

    s = prob.spec
    theta = prob.theta0()
    rng = np.random.default_rng(seed)
    dw = _directions(prob, theta, rng)
    sim = prob.build(theta, precision)
    phi_a, grads = sim.adjoint_run(s.steps, s.checkpoint_every, s.window_steps, *s.window)
    sim.close()
    sim = prob.build(theta, precision)
    phi_t, dphi = sim.tangent_run(s.steps, s.window_steps, *s.window, dw)
    sim.close()
    adj = np.array([sum(float(np.dot(g, d[k])) for g, d in zip(grads, dw)) for k in range(capi.TANGENT_DIRS)])
    rel = np.abs(adj - dphi) / np.abs(dphi)
    print("phi adjoint run", phi_a, "tangent run", phi_t)
    print("<grad, v> (adjoint)   ", np.array2string(adj, precision=12))
    print("d phi along v (tangent)", np.array2string(dphi, precision=12))
    print("relative difference    ", np.array2string(rel, precision=3))
    return phi_a, phi_t, rel


def test_dot_product_reduced_fp64():
    from pdp3_amd import adjoint as adj
    from pdp3_amd import capi

    print("dot-product test, reduced problem (32x256, 200 steps), fp64", flush=True)
    spec = adj.cfg5_spec(cells_r=32, cells_z=256, steps=200, checkpoint_every=30, window_steps=60, ppc=4,
                         bunch_macro=4000, window_frac=(0.2, 0.4))
    phi_a, phi_t, rel = _dot_product(adj.BumpRampProblem(spec), capi.PREC_FP64, 1)
    assert abs(phi_a - phi_t) <= 1e-12 * abs(phi_a)
    assert np.all(rel <= 1e-10), rel


def test_dot_product_reduced_fp32():
    from pdp3_amd import adjoint as adj
    from pdp3_amd import capi

    print("dot-product test, reduced problem (32x256, 200 steps), fp32 forward", flush=True)
    spec = adj.cfg5_spec(cells_r=32, cells_z=256, steps=200, checkpoint_every=30, window_steps=60, ppc=4,
                         bunch_macro=4000, window_frac=(0.2, 0.4))
    phi_a, phi_t, rel = _dot_product(adj.BumpRampProblem(spec), capi.PREC_FP32, 2)
    assert abs(phi_a - phi_t) <= 1e-5 * abs(phi_a)  # two FP32 forward runs: deposit sums in FP64 atomics
    assert np.all(rel <= 1e-8), rel


def _dot_product_one_sweep(prob, precision, seed, checkpoint_every):
    from pdp3_amd import capi

    s = prob.spec
    theta = prob.theta0()
    dw = _directions(prob, theta, np.random.default_rng(seed))
    sim = prob.build(theta, precision)
    phi, grads, dphi = sim.adjoint_tangent_run(s.steps, checkpoint_every, s.window_steps, *s.window, dw)
    sim.close()
    adj = np.array([sum(float(np.dot(g, d[k])) for g, d in zip(grads, dw)) for k in range(capi.TANGENT_DIRS)])
    rel = np.abs(adj - dphi) / np.abs(dphi)
    print("checkpoint every", checkpoint_every, "phi", phi)
    print("<grad, v> (adjoint)   ", np.array2string(adj, precision=12))
    print("d phi along v (tangent)", np.array2string(dphi, precision=12))
    print("relative difference    ", np.array2string(rel, precision=3), flush=True)
    return rel


@pytest.mark.parametrize("checkpoint_every,bar", [(1, 1e-8), (50, 1e-4)])
def test_dot_product_cfg5_fp64(checkpoint_every, bar):
    from pdp3_amd import adjoint as adj
    from pdp3_amd import capi

    print("dot-product test, cfg5 (128x1024, 1500 steps), fp64, one forward sweep, checkpoints every",
          checkpoint_every, flush=True)
    rel = _dot_product_one_sweep(adj.BumpRampProblem(adj.cfg5_spec()), capi.PREC_FP64, 3, checkpoint_every)
    assert np.all(rel <= bar), rel


def test_fd_converges_to_adjoint_smooth(orc):
    """On the smooth reduced problem the oracle's central differences approach
    the adjoint as the step shrinks: on every dominant parameter the error at
    a relative step of 1e-4 is within 1e-4 of the adjoint and at most a fifth
    of the largest error seen at the larger steps (3e-2, 3e-3)."""
    from pdp3_amd import adjoint as adj
    from pdp3_amd import capi

    from test_gpu_adjoint import oracle_objective

    spec = adj.cfg5_spec(cells_r=32, cells_z=256, steps=200, checkpoint_every=30, window_steps=60, ppc=4,
                         bunch_macro=4000, window_frac=(0.2, 0.4), plasma_margin_cells=3, bunch_r_min=5e-4,
                         bunch_vr_spread=1e4)
    prob = adj.BumpRampProblem(spec)
    th = prob.theta0()
    theta = th.clone().requires_grad_(True)
    phi = adj.PicObjective.apply(theta, prob, capi.PREC_FP64)
    phi.backward()
    g = theta.grad.numpy()
    dom = np.argsort(-np.abs(g))[:3]
    print("adjoint", np.array2string(g, precision=9), "dominant", dom, flush=True)
    for j in dom:
        errs = []
        for h in (3e-2, 3e-3, 1e-4):
            step = h * abs(float(th[j]))
            tp, tm = th.clone(), th.clone()
            tp[j] += step
            tm[j] -= step
            fd = (oracle_objective(orc, prob, tp) - oracle_objective(orc, prob, tm)) / (2 * step)
            errs.append(abs(fd - g[j]) / abs(g[j]))
        print("parameter", j, "relative |fd - adjoint| at h = 3e-2, 3e-3, 1e-4:", errs, flush=True)
        assert errs[2] <= 1e-4, (j, errs)
        assert errs[2] <= 0.2 * max(errs[0], errs[1]), (j, errs)

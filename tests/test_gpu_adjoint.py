"""Adjoint gradients against central finite differences.

A reduced copy of the cfg5 problem (32 x 256 cells, 200 steps, checkpoints
every 30 steps so that the last segment is partial, 16 bump amplitudes and
the ramp slope): the gradient of the axis E_z^2 objective from the library's
adjoint, chained through torch (PicObjective), against central differences
(relative step 1e-4) of the FP64 plain form, whose loop is the oracle-pinned
reference arithmetic.  FP64 forward: 1e-4 relative on every component above 1%
of the largest; FP32 forward (the production form): 2e-2 relative, its
forward error.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _problem():
    from pdp3_amd import adjoint as adj

    spec = adj.cfg5_spec(cells_r=32, cells_z=256, steps=200, checkpoint_every=30, window_steps=60, ppc=4,
                         bunch_macro=4000, window_frac=(0.2, 0.4))
    return adj.BumpRampProblem(spec)


@pytest.fixture(scope="module")
def problem_and_fd():
    from pdp3_amd import adjoint as adj

    prob = _problem()
    th = prob.theta0()
    print("adjoint: central differences of the FP64 form, 17 parameters", flush=True)
    fd = adj.central_fd(prob, th)
    return prob, th, fd


def _check(g, fd, tol):
    big = np.abs(fd) > 1e-2 * np.max(np.abs(fd))
    rel = np.abs(g[big] - fd[big]) / np.abs(fd[big])
    print("fd", fd)
    print("adjoint", g)
    print("max rel", rel.max(), "over", big.sum(), "components")
    assert big.sum() >= 5
    assert rel.max() <= tol, rel


def test_adjoint_fp64_matches_fd(problem_and_fd):
    from pdp3_amd import adjoint as adj
    from pdp3_amd import capi

    prob, th, fd = problem_and_fd
    theta = th.clone().requires_grad_(True)
    phi = adj.PicObjective.apply(theta, prob, capi.PREC_FP64)
    phi.backward()
    phi_fwd = prob.objective(th)
    assert abs(float(phi) - phi_fwd) <= 1e-9 * abs(phi_fwd)
    _check(theta.grad.numpy(), fd, 1e-4)


def test_adjoint_fp32_matches_fd(problem_and_fd):
    from pdp3_amd import adjoint as adj
    from pdp3_amd import capi

    prob, th, fd = problem_and_fd
    theta = th.clone().requires_grad_(True)
    phi = adj.PicObjective.apply(theta, prob, capi.PREC_FP32)
    phi.backward()
    _check(theta.grad.numpy(), fd, 2e-2)

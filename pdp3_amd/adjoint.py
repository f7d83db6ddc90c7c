"""Gradients of a probe E_z-energy objective with respect to bunch-profile and
density-ramp parameters, through the hand-written adjoint of the C ABI
(pdp3_adjoint_run), wrapped in a torch.autograd.Function.

The parameters enter the simulation through particle weights only (real
particles per macro-particle): the plasma weight follows a linear density
ramp whose slope is a parameter, and the bunch weight an axial profile that is
a sum of Gaussian bumps whose amplitudes are parameters.  Positions and
velocities of the initial state do not depend on them.  The library returns
dphi/dw for every particle; torch chains it through w(theta).

    problem = BumpRampProblem(cfg5_spec())
    theta = problem.theta0().requires_grad_()
    phi = PicObjective.apply(theta, problem)   # forward: FP32 form + adjoint
    phi.backward()                              # theta.grad
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import capi
from .workloads import EL_MASS, LIGHT_SPEED, PI, DEFAULT_SEED


@dataclass
class AdjointSpec:
    """The adjoint configuration: grid, plasma with a linear ramp, a bunch
    whose axial profile is a sum of Gaussian bumps, and the objective (E_z^2
    on the axis over [z_lo, z_hi] across the last window_steps steps)."""
    cells_r: int = 128
    cells_z: int = 1024
    dr: float = 5e-4
    steps: int = 1500
    checkpoint_every: int = 50
    window_steps: int = 300
    ppc: int = 20
    density: float = 1e17
    slope: float = 0.5          # n(z) = density * (1 + slope * (z / L - 1/2))
    n_bumps: int = 16
    bump_sigma: float = 1e-3
    bump_span: float = 16e-3    # centres evenly spaced over this length
    bump_start: float = 2e-3    # centre of the first bump
    amp_max: float = 2e16       # amplitudes uniform in 0 .. amp_max
    bunch_radius: float = 1.5e-3
    bunch_macro: int = 64000
    velocity: float = 0.99 * LIGHT_SPEED
    temperature_ev: float = 1.0
    pml_cells: int = 8
    window_frac: tuple = (0.7, 0.9)
    # plasma particles closer than this many cells to the axis or a wall are
    # left out (0: none), and the smallest bunch radius: with a cold plasma
    # they keep every particle away from the reflecting boundaries
    plasma_margin_cells: int = 0
    bunch_r_min: float = 0.0
    # spread (m/s, uniform in +-) of the bunch's initial radial velocity; 0
    # starts every bunch particle on a purely axial path
    bunch_vr_spread: float = 0.0
    seed: int = DEFAULT_SEED

    @property
    def length(self):
        return self.cells_z * self.dr

    @property
    def window(self):
        return self.window_frac[0] * self.length, self.window_frac[1] * self.length


def cfg5_spec(**over) -> AdjointSpec:
    """cfg5: 128 x 1024 cells, dr = dz = 5e-4 m, 1500 steps, 20 ppc at 1e17 with
    a ramp slope parameter, 16 bumps (sigma 1 mm over 16 mm, amplitudes in
    0..2e16, radius 1.5 mm, 0.99 c), E_z^2 on axis over [0.7 L, 0.9 L] in the
    last 300 steps; checkpoints every 50 steps."""
    s = AdjointSpec()
    for k, v in over.items():
        setattr(s, k, v)
    return s


class BumpRampProblem:
    """Fixed initial positions / velocities; weights as differentiable
    functions of theta = (16 bump amplitudes in units of amp_max, ramp slope)."""

    def __init__(self, spec: AdjointSpec):
        self.spec = spec
        s = spec
        self.n_grid_r, self.n_grid_z = s.cells_r + 1, s.cells_z + 1
        self.dt = 0.5 * s.dr / LIGHT_SPEED
        rng = np.random.default_rng(s.seed + 5)
        self.amp0 = rng.uniform(0.0, 1.0, s.n_bumps)
        # plasma: the device loader's uniform state (weights of the flat density)
        sim = self._new_sim(capi.PREC_FP64)
        st = sim.download(0)
        sim.close()
        if s.plasma_margin_cells > 0:
            m = s.plasma_margin_cells * s.dr
            R, L = s.cells_r * s.dr, s.cells_z * s.dr
            out = (st["r"] < m) | (st["r"] > R - m) | (st["z"] < m) | (st["z"] > L - m)
            st["alive"] = np.where(out, 0, st["alive"]).astype(np.uint8)
        self.plasma = st
        self.plasma_z = torch.tensor(st["z"], dtype=torch.float64)
        self.plasma_w0 = torch.tensor(st["w"], dtype=torch.float64)
        # bunch: radii uniform in area within the radius, axial positions
        # uniform over the bump train, at the bunch velocity along z
        n = s.bunch_macro
        z_a = max(s.bump_start - 3 * s.bump_sigma, s.dr)
        z_b = s.bump_start + s.bump_span + 3 * s.bump_sigma
        self.bunch_r = s.bunch_radius * np.sqrt(rng.uniform(0.0, 1.0, n))
        self.bunch_r = np.maximum(self.bunch_r, max(s.dr / 4, s.bunch_r_min))
        self.bunch_z = rng.uniform(z_a, z_b, n)
        self.bunch_vr = rng.uniform(-1.0, 1.0, n) * s.bunch_vr_spread
        # the same samples in order of z: the bunch kernel gathers and deposits
        # a block's particles through an LDS box anchored at their smallest
        # cell, which only helps when a block's particles lie close together
        order = np.argsort(self.bunch_z, kind="stable")
        self.bunch_r, self.bunch_z, self.bunch_vr = self.bunch_r[order], self.bunch_z[order], self.bunch_vr[order]
        # each macro-particle carries the volume pi R^2 (z_b - z_a) / n
        self.bunch_vol = PI * s.bunch_radius ** 2 * (z_b - z_a) / n
        centres = s.bump_start + np.arange(s.n_bumps) * (s.bump_span / max(s.n_bumps - 1, 1))
        prof = np.exp(-(self.bunch_z[:, None] - centres[None, :]) ** 2 / (2 * s.bump_sigma ** 2))
        self.bunch_basis = torch.tensor(prof * s.amp_max * self.bunch_vol, dtype=torch.float64)

    def _new_sim(self, precision):
        s = self.spec
        sim = capi.Simulation(self.n_grid_r, self.n_grid_z, s.cells_r * s.dr, s.cells_z * s.dr, self.dt,
                              precision=precision)
        if s.pml_cells:
            n = s.pml_cells
            sim.set_pml((n + 0.5) / self.n_grid_z, (n + 0.5) / self.n_grid_z, (n + 0.5) / self.n_grid_r, 1e-5, 7e-2)
        sim.boundary_conditions()
        idx = sim.add_species(-1, 1.0, s.ppc * s.cells_r * s.cells_z)
        sim.init_uniform(idx, s.seed, s.ppc, s.density, 0, s.temperature_ev)
        return sim

    def theta0(self):
        return torch.tensor(np.concatenate([self.amp0, [self.spec.slope]]), dtype=torch.float64)

    def weights(self, theta):
        """[plasma weights, bunch weights] as torch functions of theta"""
        L = self.spec.length
        slope = theta[-1]
        w_plasma = self.plasma_w0 * (1 + slope * (self.plasma_z / L - 0.5))
        w_bunch = self.bunch_basis @ theta[:-1]
        return [w_plasma, w_bunch]

    def build(self, theta, precision):
        """a Simulation holding the initial state for theta"""
        s = self.spec
        w_plasma, w_bunch = (w.detach().cpu().numpy() for w in self.weights(theta))
        sim = self._new_sim(precision)
        sim.upload(0, w=w_plasma, alive=self.plasma["alive"])
        n = len(self.bunch_r)
        zero = np.zeros(n)
        sim.add_bunch_loaded(-1, 1.0, self.bunch_r, self.bunch_z, self.bunch_vr, zero, np.full(n, s.velocity), w_bunch)
        return sim

    def objective(self, theta, precision=capi.PREC_FP64):
        """phi(theta), forward only"""
        s = self.spec
        sim = self.build(theta, precision)
        try:
            return sim.objective_run(s.steps, s.window_steps, *s.window)
        finally:
            sim.close()

    def objective_and_grad_w(self, theta, precision=capi.PREC_FP32):
        s = self.spec
        sim = self.build(theta, precision)
        try:
            return sim.adjoint_run(s.steps, s.checkpoint_every, s.window_steps, *s.window)
        finally:
            sim.close()


    def directional_derivatives(self, theta, dthetas, precision=capi.PREC_FP64):
        """phi(theta) and its derivatives along up to TANGENT_DIRS parameter
        directions (rows of dthetas) in one forward-mode run of the library's
        tangent-linear model (pdp3_tangent_run)"""
        s = self.spec
        dthetas = torch.as_tensor(dthetas, dtype=torch.float64).reshape(-1, len(theta))
        k = dthetas.shape[0]
        if k > capi.TANGENT_DIRS:
            raise ValueError(f"at most {capi.TANGENT_DIRS} directions per run")
        th = torch.as_tensor(theta, dtype=torch.float64).detach()
        dws = [np.zeros((capi.TANGENT_DIRS, n)) for n in (len(self.plasma["r"]), len(self.bunch_r))]
        for d in range(k):
            _, tangents = torch.func.jvp(lambda t: tuple(self.weights(t)), (th,), (dthetas[d],))
            for j, t in enumerate(tangents):
                dws[j][d] = t.numpy()
        sim = self.build(th, precision)
        try:
            phi, dphi = sim.tangent_run(s.steps, s.window_steps, *s.window, dws)
        finally:
            sim.close()
        return phi, dphi[:k]


    def batch_gradients(self, thetas, precision=capi.PREC_FP32):
        """phi and dphi/dtheta at several parameter points in one batched
        adjoint run (pdp3_adjoint_run_batch: the members run concurrently on
        their own streams): (phis [m], grads [m, len(theta)])"""
        s = self.spec
        thetas = [torch.as_tensor(t, dtype=torch.float64).detach() for t in thetas]
        sims = [self.build(t, precision) for t in thetas]
        try:
            phis, grads_w = capi.adjoint_run_batch(sims, s.steps, s.checkpoint_every, s.window_steps, *s.window)
        finally:
            for sim in sims:
                sim.close()
        out = []
        for t, gw in zip(thetas, grads_w):
            th = t.clone().requires_grad_(True)
            ws = self.weights(th)
            (g,) = torch.autograd.grad(ws, th, grad_outputs=[torch.tensor(a, dtype=torch.float64) for a in gw])
            out.append(g)
        return np.array(phis), torch.stack(out).numpy()


class PicObjective(torch.autograd.Function):
    """phi = PicObjective.apply(theta, problem[, precision]): the objective of
    the r-z PIC loop; backward through the library's adjoint."""

    @staticmethod
    def forward(ctx, theta, problem, precision=capi.PREC_FP32):
        phi, grads = problem.objective_and_grad_w(theta.detach(), precision)
        ctx.save_for_backward(theta)
        ctx.problem = problem
        ctx.grads = [torch.tensor(g, dtype=torch.float64) for g in grads]
        return theta.new_tensor(phi)

    @staticmethod
    def backward(ctx, grad_out):
        (theta,) = ctx.saved_tensors
        with torch.enable_grad():
            th = theta.detach().requires_grad_(True)
            ws = ctx.problem.weights(th)
            (g,) = torch.autograd.grad(ws, th, grad_outputs=ctx.grads[:len(ws)])
        return grad_out * g, None, None


def central_fd(problem, theta, rel_h=1e-4, precision=capi.PREC_FP64, which=None):
    """central finite differences of phi (FP64 plain form by default), step
    rel_h |theta_j| (rel_h for a zero entry)"""
    theta = theta.detach().clone()
    out = np.zeros(len(theta))
    for j in (range(len(theta)) if which is None else which):
        h = rel_h * max(abs(float(theta[j])), 1.0 if theta[j] == 0 else 0.0)
        tp, tm = theta.clone(), theta.clone()
        tp[j] += h
        tm[j] -= h
        out[j] = (problem.objective(tp, precision) - problem.objective(tm, precision)) / (2 * h)
    return out

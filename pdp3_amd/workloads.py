"""Synthetic workloads of the benchmark configurations (r-z grids, plasma,
bunches), built on top of the C ABI.

Plasma comes from the device loader (pdp3_species_init_uniform: ppc particles
per cell, uniform in-cell positions, isotropic Maxwellian, Philox draws keyed
on the seed).  Bunches are described by (r, w, inject_u) arrays as the
reference's Bunch constructor leaves them (src/bunch.cpp:6-80) and injected
over time by Bunch::bunch_inject (src/bunch.cpp:87-124); they are drawn here
with numpy from the same seed.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

LIGHT_SPEED = 3.0e8
EL_CHARGE = 1.6e-19
EL_MASS = 9.1e-31
PI = 3.1415926535897932
DEFAULT_SEED = 20261202


@dataclass
class Bunch:
    charge: int
    mass: float
    r: np.ndarray
    w: np.ndarray
    inject_u: np.ndarray
    duration: float
    hole_duration: float
    velocity: float
    bunch_number: int = 0


@dataclass
class Plasma:
    charge: int
    mass: float
    ppc: int
    density: float
    temperature_ev: float
    ramp_cells: int = 0


@dataclass
class Workload:
    name: str
    cells_r: int
    cells_z: int
    dr: float
    dz: float
    steps: int
    pml_cells: int
    pusher: int  # 0 Boris, 1 Vay
    plasma: list = field(default_factory=list)
    bunches: list = field(default_factory=list)
    seed: int = DEFAULT_SEED

    @property
    def n_grid_r(self):
        return self.cells_r + 1

    @property
    def n_grid_z(self):
        return self.cells_z + 1

    @property
    def r_size(self):
        return self.cells_r * self.dr

    @property
    def z_size(self):
        return self.cells_z * self.dz

    @property
    def dt(self):
        return 0.5 * self.dr / LIGHT_SPEED

    def pml(self):
        """set_pml arguments giving exactly pml_cells cells (src/geometry.cpp:60-66:
        L = dz * floor(n_grid * l))."""
        n = self.pml_cells
        if n == 0:
            return None
        return ((n + 0.5) / self.n_grid_z, (n + 0.5) / self.n_grid_z, (n + 0.5) / self.n_grid_r, 1e-5, 7e-2)

    def plasma_count(self, p: Plasma):
        return p.ppc * self.cells_r * self.cells_z

    def particle_count(self):
        return sum(self.plasma_count(p) for p in self.plasma) + sum(len(b.r) for b in self.bunches)


def uniform_bunch(rng, n, radius, length, density, velocity, dr, dz, charge=-1, mass=1.0):
    """The reference's uniform bunch (src/bunch.cpp:6-80): radii uniform in
    [dr/2, radius - dr/2], weights proportional to r normalised to
    pi radius^2 length density real particles; injection draws uniform."""
    duration = length / velocity
    r = (radius - dr) * rng.random(n) + dr / 2
    vol = np.where(r > dr, 2 * PI * r * dr * dz, PI * r * r * dz)
    v_avg = vol.sum() / n
    n_total = PI * radius ** 2 * length * density
    norm = np.where(r > dr / 2, 2 * PI * r * dr * dz / v_avg, PI * (r * r + r * dr + dr * dr / 4) * dz / v_avg)
    w = n_total / n * norm
    u = rng.random(n)
    return Bunch(charge, mass, r, w, u, duration, 0.0, velocity)


def gaussian_bunch(rng, n, sigma_r, sigma_z, peak_density, velocity, dr, dz, charge=-1, mass=1.0):
    """A Gaussian bunch (sigma_r, sigma_z, peak density) expressed in the
    reference's injection model: the bunch enters at z = dz/2 over
    duration = 6 sigma_z / v, particle j in step j // (n / steps); radii
    uniform in [dr/2, 3 sigma_r] and weights carrying the density
    n_peak exp(-r^2/2 sigma_r^2) exp(-(z-3 sigma_z)^2/2 sigma_z^2), z the
    particle's position along the bunch (j / n * 6 sigma_z)."""
    length = 6.0 * sigma_z
    duration = length / velocity
    r_max = max(3.0 * sigma_r, dr)
    r = (r_max - dr / 2) * rng.random(n) + dr / 2
    zeta = (np.arange(n) + 0.5) / n * length
    dens = peak_density * np.exp(-r ** 2 / (2 * sigma_r ** 2)) * np.exp(-(zeta - 3 * sigma_z) ** 2 / (2 * sigma_z ** 2))
    # each macro-particle stands for the ring slice 2 pi r (r_max - dr/2) (length / n) of the profile
    w = dens * 2 * PI * r * (r_max - dr / 2) * (length / n)
    u = rng.random(n)
    return Bunch(charge, mass, r, w, u, duration, 0.0, velocity)


def config(name: str, seed: int = DEFAULT_SEED, bunch_macro: int | None = None, ramp_cells: int | None = None,
           **over) -> Workload:
    """The benchmark configurations.  cfg1: 64x256 CPU-sized case; cfg2:
    256x2048 production size with mobile ions and Vay; cfg3: 512x8192 at
    100 ppc with a density ramp (the precision-form comparison).
    bunch_macro and ramp_cells override the bunch's macro-particle count and
    the density ramp's length in z-cells (reduced copies for parity tests);
    other keywords override Workload fields (cells_r, cells_z, pml_cells,
    pusher, steps)."""
    rng = np.random.default_rng(seed)
    if name == "cfg1":
        dr = dz = 5e-4
        w = Workload("cfg1", 64, 256, dr, dz, 500, 8, 0, seed=seed)
        w.plasma = [Plasma(-1, 1.0, 10, 1e17, 1.0)]
        w.bunches = [uniform_bunch(rng, bunch_macro or 1000, 2e-3, 5e-3, 1e16, 0.9 * LIGHT_SPEED, dr, dz)]
    elif name == "cfg2":
        dr = dz = 2.5e-4
        w = Workload("cfg2", 256, 2048, dr, dz, 2000, 16, 1, seed=seed)
        w.plasma = [Plasma(-1, 1.0, 40, 1e17, 1.0), Plasma(1, 1836.0, 40, 1e17, 0.1)]
        w.bunches = [gaussian_bunch(rng, bunch_macro or 100000, 1e-3, 2e-3, 2e16, 0.99 * LIGHT_SPEED, dr, dz)]
    elif name == "cfg3":
        dr = dz = 2.5e-4
        w = Workload("cfg3", 512, 8192, dr, dz, 1000, 16, 0, seed=seed)
        w.plasma = [Plasma(-1, 1.0, 100, 1e18, 1.0, ramp_cells=1000 if ramp_cells is None else ramp_cells)]
        w.bunches = [gaussian_bunch(rng, bunch_macro or 1000000, 0.5e-3, 1e-3, 5e17, 0.999 * LIGHT_SPEED, dr, dz)]
    else:
        raise ValueError(name)
    for k, v in over.items():
        setattr(w, k, v)
    return w


def build_simulation(w: Workload, precision: int = 0, sort_interval: int = 0):
    """A Simulation holding the workload's initial state on the GPU."""
    from .capi import Simulation

    sim = Simulation(w.n_grid_r, w.n_grid_z, w.r_size, w.z_size, w.dt, pusher=w.pusher, precision=precision)
    pml = w.pml()
    if pml:
        sim.set_pml(*pml)
    sim.boundary_conditions()
    for j, p in enumerate(w.plasma):
        idx = sim.add_species(p.charge, p.mass, w.plasma_count(p))
        sim.init_uniform(idx, w.seed + 7919 * j, p.ppc, p.density, p.ramp_cells, p.temperature_ev)
    for b in w.bunches:
        sim.add_bunch(b.charge, b.mass, b.r, b.w, b.inject_u, b.duration, b.hole_duration, b.velocity,
                      b.bunch_number)
    if sort_interval:
        sim.set_sort_interval(sort_interval)
    return sim


@dataclass
class Instance:
    """One member of the cfg4 ensemble."""
    cells_r: int
    cells_z: int
    ppc: int
    density: float
    bunch_ratio: float
    seed: int
    dr: float = 5e-4
    pml_cells: int = 8
    temperature_ev: float = 1.0
    bunch_macro: int = 20000
    sigma_r: float = 1e-3
    sigma_z: float = 2e-3
    bunch_velocity: float = 0.99 * LIGHT_SPEED

    @property
    def particles(self):
        return self.ppc * self.cells_r * self.cells_z + self.bunch_macro


def ensemble_config(n_instances: int = 128, seed: int = DEFAULT_SEED):
    """cfg4: per-instance draws from the seed: r cells in {64, 128, 256}, z
    cells in {512, 1024, 2048, 4096}, dr = dz = 5e-4 m, electron density
    log-uniform in 1e16..1e18 m^-3, 5..40 ppc, 1 eV, bunch / plasma density
    ratio uniform in 0.01..0.5, Gaussian bunch (sigma_r 1 mm, sigma_z 2 mm) at
    0.99 c."""
    rng = np.random.default_rng(seed + 4)
    out = []
    for j in range(n_instances):
        out.append(Instance(cells_r=int(rng.choice([64, 128, 256])), cells_z=int(rng.choice([512, 1024, 2048, 4096])),
                            ppc=int(rng.integers(5, 41)), density=float(10 ** rng.uniform(16, 18)),
                            bunch_ratio=float(rng.uniform(0.01, 0.5)), seed=seed + 1009 * j))
    return out


def loaded_gaussian_bunch(rng, inst: Instance):
    """The Gaussian bunch already inside the grid: its 6 sigma_z window starts
    at z = dz/2, every particle moves at the bunch velocity along z."""
    dr = dz = inst.dr
    b = gaussian_bunch(rng, inst.bunch_macro, inst.sigma_r, inst.sigma_z, inst.bunch_ratio * inst.density,
                       inst.bunch_velocity, dr, dz)
    n = inst.bunch_macro
    z = (np.arange(n) + 0.5) / n * (6.0 * inst.sigma_z) + dz / 2
    zero = np.zeros(n)
    return b.r, z, zero, zero, np.full(n, inst.bunch_velocity), b.w


def build_instance(inst: Instance, precision: int = 1, injected: bool = False):
    """A Simulation of one ensemble member (FP32 form by default).  The
    Gaussian bunch is loaded inside the grid at t = 0, or, with injected,
    enters through z = dz/2 over 6 sigma_z / v as the reference's Bunch does
    (Bunch::bunch_inject, src/bunch.cpp:87-124)."""
    from .capi import Simulation

    w = Workload("cfg4", inst.cells_r, inst.cells_z, inst.dr, inst.dr, 1000, inst.pml_cells, 0, seed=inst.seed)
    sim = Simulation(w.n_grid_r, w.n_grid_z, w.r_size, w.z_size, w.dt, pusher=0, precision=precision)
    pml = w.pml()
    if pml:
        sim.set_pml(*pml)
    sim.boundary_conditions()
    idx = sim.add_species(-1, 1.0, inst.ppc * inst.cells_r * inst.cells_z)
    sim.init_uniform(idx, inst.seed, inst.ppc, inst.density, 0, inst.temperature_ev)
    rng = np.random.default_rng(inst.seed + 1)
    if injected:
        b = gaussian_bunch(rng, inst.bunch_macro, inst.sigma_r, inst.sigma_z, inst.bunch_ratio * inst.density,
                           inst.bunch_velocity, w.dr, w.dz)
        sim.add_bunch(b.charge, b.mass, b.r, b.w, b.inject_u, b.duration, b.hole_duration, b.velocity,
                      b.bunch_number)
    else:
        sim.add_bunch_loaded(-1, 1.0, *loaded_gaussian_bunch(rng, inst))
    return sim


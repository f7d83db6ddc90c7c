"""GPU parity of the FP64 plain form against the oracle, through the C ABI.

Kernels without atomics (field updates, gathers, per-particle push) must be
bit-identical to the oracle: both evaluate every expression in the
reference's order without FMA contraction.  Grid sums are order-dependent
(hardware FP64 atomics vs the reference's serial loop), so J / rho and
anything downstream are compared normwise: <= 1e-13 after one deposit and,
after 100 steps:

* <= 1e-10, strictly, on the reduced cfg2 in test_gpu_configs.py;
* on the 64x256 configuration the reference's loop itself moves by more than
  1e-10 when only the order of its serial deposit sums changes, or when only
  the last bit of lib::get_gamma_inv changes in the few hundredths of a
  percent of arguments where glibc's pow(y, -0.5) is not the double nearest
  to 1/sqrt(y) (the device computes the nearest): with Boris on the plasma
  alone H_phi moves by 2.5e-10 and J by 4.4e-9 (a particle reflected at the
  axis one step earlier or later), with the bunch E_z by ~7e-9; with Vay J_r
  moves by 4.0e-10 on the plasma alone and by 1.2e-10 with the bunch, from
  the same kind of event.  The GPU's deposit sums (FP64 atomics, in an order
  that varies from run to run) land on either side of such an event: one run
  of the Vay plasma agreed with the oracle to 1e-10, another matched the
  oracle's perturbed run to 6 digits (J_r 4.00061e-10 against 4.00061e-10).
  The bar is therefore 10x the largest such self-change of the oracle, and
  never more than a fixed ceiling (2e-8 for fields, 5e-8 for particle state)
  set from those recorded floors.
"""
import numpy as np
import pytest

from helpers import compare_fields, compare_species, normwise, oracle_from_workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    from pdp3_amd import capi as c
    c.lib()
    return c


def _small_geometry():
    return dict(n_grid_r=33, n_grid_z=129, r_size=32 * 5e-4, z_size=128 * 5e-4)


# node counts: the small geometry, and grids whose sizes leave partial field
# tiles in both directions with several tiles along z (tiles of 2 x 256 nodes
# for H, 1 x 256 for E)
@pytest.mark.parametrize("nr,nz", [(33, 129), (34, 600), (7, 1030)])
def test_field_updates_bitwise(capi, orc, nr, nz):
    g = dict(n_grid_r=nr, n_grid_z=nz, r_size=(nr - 1) * 5e-4, z_size=(nz - 1) * 5e-4)
    dt = 0.5 * 5e-4 / 3e8
    rng = np.random.default_rng(1)
    gpu = capi.Simulation(g["n_grid_r"], g["n_grid_z"], g["r_size"], g["z_size"], dt)
    ora = orc.Sim(g["n_grid_r"], g["n_grid_z"], g["r_size"], g["z_size"], dt)
    pml = ((6.5) / nz, (6.5) / nz, (min(4, nr - 3) + 0.5) / nr, 1e-5, 7e-2)
    gpu.set_pml(*pml)
    ora.set_pml(*pml)
    assert np.array_equal(gpu.field("sigma"), ora.field("sigma"))
    for name in ("E_r", "E_phi", "E_z", "H_r", "H_phi", "H_z", "J_r", "J_phi", "J_z"):
        a = rng.standard_normal((g["n_grid_r"], g["n_grid_z"])) * (1e3 if name[0] == "E" else 1.0)
        gpu.set_field(name, a)
        ora.field(name)[:] = a
    for _ in range(3):
        gpu.h_calc()
        orc.lib().or_sim_h_calc(ora.h)
        gpu.e_calc()
        orc.lib().or_sim_e_calc(ora.h)
    for name in ("E_r", "E_phi", "E_z", "H_r", "H_phi", "H_z", "HH_r", "HH_phi", "HH_z"):
        assert np.array_equal(gpu.field(name), ora.field(name)), name


def test_gather_bitwise(capi, orc):
    g = _small_geometry()
    dt = 0.5 * 5e-4 / 3e8
    rng = np.random.default_rng(2)
    gpu = capi.Simulation(g["n_grid_r"], g["n_grid_z"], g["r_size"], g["z_size"], dt)
    ora = orc.Sim(g["n_grid_r"], g["n_grid_z"], g["r_size"], g["z_size"], dt)
    for name in ("E_r", "E_phi", "E_z", "HH_r", "HH_phi", "HH_z"):
        a = rng.standard_normal((g["n_grid_r"], g["n_grid_z"]))
        gpu.set_field(name, a)
        ora.field(name)[:] = a
    dr, dz = gpu.dr, gpu.dz
    n = 4000
    r = rng.uniform(dr / 2, g["r_size"] - dr / 2, n)
    z = rng.uniform(dz / 2, g["z_size"] - dz / 2, n)
    # near the axis, on nodes, on half nodes and at the walls
    r[:200] = rng.uniform(dr / 2, 1.5 * dr, 200)
    r[200:260] = np.arange(1, 61) % 30 * dr + dr
    z[260:320] = (np.arange(60) % 120 + 1) * dz
    r[320:380] = (np.arange(60) % 30 + 0.5) * dr + dr / 2
    r[380] = dr / 2
    z[381] = dz / 2
    r[382] = g["r_size"] - dr / 2
    z[383] = g["z_size"] - dz / 2
    ge, gh = gpu.get_e(r, z), gpu.get_h(r, z)
    oe = np.array([ora.get_e(a, b) for a, b in zip(r, z)])
    oh = np.array([ora.get_h(a, b) for a, b in zip(r, z)])
    assert np.array_equal(ge, oe)
    assert np.array_equal(gh, oh)


def _random_particles(rng, n, g, dr, dz, vscale):
    r = rng.uniform(dr / 2, g["r_size"] - dr / 2, n)
    z = rng.uniform(dz / 2, g["z_size"] - dz / 2, n)
    v = rng.standard_normal((3, n)) * vscale
    speed = np.sqrt((v ** 2).sum(axis=0))
    v *= np.minimum(1.0, 0.8 * 3e8 / speed)  # keep |v| < 0.8 c
    w = rng.uniform(1e5, 2e5, n)
    return r, z, v[0], v[1], v[2], w


@pytest.mark.parametrize("pusher", [0, 1])
def test_split_step_matches_oracle(capi, orc, pusher):
    """Particles::step_v then full_current_distribution through the split
    entry points: particle state bit-identical, J normwise <= 1e-13."""
    g = _small_geometry()
    dt = 0.5 * 5e-4 / 3e8
    rng = np.random.default_rng(3 + pusher)
    gpu = capi.Simulation(g["n_grid_r"], g["n_grid_z"], g["r_size"], g["z_size"], dt, pusher=pusher)
    ora = orc.Sim(g["n_grid_r"], g["n_grid_z"], g["r_size"], g["z_size"], dt, pusher=pusher)
    for name in ("E_r", "E_phi", "E_z", "HH_r", "HH_phi", "HH_z"):
        a = rng.standard_normal((g["n_grid_r"], g["n_grid_z"])) * (2e4 if name[0] == "E" else 20.0)
        gpu.set_field(name, a)
        ora.field(name)[:] = a
    # fast particles so that many cross cells, and a share moving strictly
    # axially / radially to reach strict_motion_weighting
    r, z, vr, vphi, vz, w = _random_particles(rng, 20000, g, gpu.dr, gpu.dz, 6e7)
    vr[:1000] = 0.0
    vz[1000:2000] = 0.0
    vphi[:2000] = 0.0
    ig = gpu.add_plasma(-1, 1.0, r, z, vr, vphi, vz, w)
    io = ora.add_plasma(-1, 1.0, r, z, vr, vphi, vz, w)
    gpu.reset_j()
    gpu.step_v(ig)
    assert orc.lib().or_sim_step_v(ora.h, io) == 0
    gpu.full_current_distribution(ig)
    assert orc.lib().or_sim_full_current_distribution(ora.h, io) == 0
    gs, os_ = gpu.download(ig), ora.get_species(io)
    diffs = {k: int(np.sum(gs[k] != os_[k])) for k in ("r", "z", "vr", "vphi", "vz")}
    jerr = {name: normwise(gpu.field(name), ora.field(name)) for name in ("J_r", "J_phi", "J_z")}
    print("mismatching particles", diffs, "J", jerr)
    # glibc's pow(y,-0.5) is not correctly rounded in ~0.04% of calls (the
    # GPU's is); such particles differ in the last bit, all others are equal
    for k in ("r", "z", "vr", "vphi", "vz"):
        assert diffs[k] <= 0.005 * len(r), (k, diffs[k])
        assert normwise(gs[k], os_[k]) <= 1e-15, k
    assert np.array_equal(gs["alive"], os_["alive"])
    for name, e in jerr.items():
        assert e <= 1e-13, (name, e)


def test_charge_weighting_matches_oracle(capi, orc):
    g = _small_geometry()
    dt = 0.5 * 5e-4 / 3e8
    rng = np.random.default_rng(5)
    gpu = capi.Simulation(g["n_grid_r"], g["n_grid_z"], g["r_size"], g["z_size"], dt)
    ora = orc.Sim(g["n_grid_r"], g["n_grid_z"], g["r_size"], g["z_size"], dt)
    r, z, vr, vphi, vz, w = _random_particles(rng, 5000, g, gpu.dr, gpu.dz, 1e5)
    r[:500] = rng.uniform(gpu.dr / 2, 1.2 * gpu.dr, 500)  # all three radial branches
    ig = gpu.add_plasma(-1, 1.0, r, z, vr, vphi, vz, w)
    io = ora.add_plasma(-1, 1.0, r, z, vr, vphi, vz, w)
    gpu.charge_weighting(ig)
    rho = ora.field("rho_bunch")
    assert orc.lib().or_sim_charge_weighting(ora.h, io, rho.ctypes.data_as(orc._D)) == 0
    assert normwise(gpu.field("rho_bunch"), rho) <= 1e-13


def _run_pair(capi, orc, w, nsteps):
    """GPU and oracle from the same initial state, plus the oracle with its
    serial deposit sums taken in other orders (reversed, and two shuffled
    particle orders) and the oracle with get_gamma_inv taken as the double
    nearest to 1/sqrt(y): how far the reference's own loop moves when only
    the order of the grid sums changes, which any parallel deposit does, or
    only the last-bit cases of glibc's pow.  Returns the GPU errors and that
    self-floor (max over the variants)."""
    from pdp3_amd import workloads as wl

    gpu = wl.build_simulation(w)
    states = [gpu.download(j) for j in range(len(w.plasma))]
    ora = oracle_from_workload(orc, w, gpu, states)
    gpu.step(nsteps)
    ora.step(nsteps)
    assert gpu.time == ora.time
    nsp = len(w.plasma) + len(w.bunches)
    errs = compare_fields(gpu, ora)
    sp = [compare_species(gpu.download(j), ora.get_species(j)) for j in range(nsp)]
    floor = {k: 0.0 for k in errs}
    sp_floor = [dict(alive_mismatch=0, r=0.0, z=0.0, vr=0.0, vphi=0.0, vz=0.0) for _ in range(nsp)]
    rng = np.random.default_rng(99)
    for variant in ("reverse", "shuffle", "shuffle", "nearest"):
        if variant == "shuffle":
            perms = [rng.permutation(len(st["r"])) for st in states]
        else:
            perms = [np.arange(len(st["r"])) for st in states]
        alt_states = [{k: v[p] for k, v in st.items()} for st, p in zip(states, perms)]
        o2 = oracle_from_workload(orc, w, gpu, alt_states)
        o2.set_reverse_deposit(variant == "reverse")
        orc.lib().or_set_gamma_inv_nearest(1 if variant == "nearest" else 0)
        try:
            o2.step(nsteps)
        finally:
            orc.lib().or_set_gamma_inv_nearest(0)
        for k, e in compare_fields(o2, ora).items():
            floor[k] = max(floor[k], e)
        for j in range(nsp):
            s2 = o2.get_species(j)
            if j < len(w.plasma):
                inv = np.argsort(perms[j])
                s2 = {k: v[inv] for k, v in s2.items()}
            for k, e in compare_species(s2, ora.get_species(j)).items():
                sp_floor[j][k] = max(sp_floor[j][k], e)
    return errs, floor, sp, sp_floor


FIELD_CEILING = 2e-8
STATE_CEILING = 5e-8


def _check(errs, floor, sp, sp_floor, strict):
    """FP64 bar: 1e-10 normwise when strict; otherwise 1e-10, or 10x the
    oracle's own self-change where the reference's loop is that sensitive,
    capped by the fixed ceilings."""
    print("fields", errs)
    print("oracle self-change floor", floor)
    print("species", sp)
    print("species self-change floor", sp_floor)
    for name, e in errs.items():
        bar = 1e-10 if strict else min(FIELD_CEILING, max(1e-10, 10 * floor[name]))
        assert e <= bar, (name, e, floor[name])
    for cs, fl in zip(sp, sp_floor):
        assert cs["alive_mismatch"] <= (0 if strict else fl["alive_mismatch"])
        for k in ("r", "z", "vr", "vphi", "vz"):
            bar = 1e-10 if strict else min(STATE_CEILING, max(1e-10, 10 * fl[k]))
            assert cs[k] <= bar, (k, cs[k], fl[k])


@pytest.mark.parametrize("pusher", [0, 1])
def test_cfg1_plasma_100_steps(capi, orc, pusher):
    """The 64x256 configuration without the bunch (10 ppc plasma at 1e17,
    1 eV, PML 8 cells), 100 full steps, within the oracle's own self-change
    (one axis reflection shifted by a step moves J by ~4e-9 with Boris, J_r
    by 4e-10 with Vay)."""
    from pdp3_amd import workloads as wl

    w = wl.config("cfg1", pusher=pusher)
    w.bunches = []
    _check(*_run_pair(capi, orc, w, 100), strict=False)


@pytest.mark.parametrize("pusher", [0, 1])
def test_cfg1_shape_100_steps(capi, orc, pusher):
    """The full 64x256 configuration (plasma + uniform 0.9c bunch, PML 8
    cells) for 100 full steps.  With Boris the reference's J_r formula
    cancels catastrophically for near-axial bunch motion, so its loop
    amplifies round-off: reordering the oracle's own serial deposit sums, or
    changing only glibc's last-bit pow cases, already moves E and J by ~1e-9
    here; with Vay a shifted axis reflection moves J_r by 1.2e-10.  Both are
    held to the oracle's own self-change (capped)."""
    from pdp3_amd import workloads as wl

    w = wl.config("cfg1", pusher=pusher)
    _check(*_run_pair(capi, orc, w, 100), strict=False)


def test_nonfinite_position_is_reported(capi):
    g = _small_geometry()
    sim = capi.Simulation(g["n_grid_r"], g["n_grid_z"], g["r_size"], g["z_size"], 1e-12)
    r = np.array([0.004, np.nan])
    z = np.array([0.01, 0.01])
    sim.add_plasma(-1, 1.0, r, z, np.zeros(2), np.zeros(2), np.zeros(2), np.ones(2))
    with pytest.raises(capi.Pdp3Error) as ei:
        sim.step(1)
    assert ei.value.code == capi.ERR_NONFINITE
    sim.sync()  # the flag is cleared after being reported


def test_superluminal_velocity_is_reported(capi):
    g = _small_geometry()
    sim = capi.Simulation(g["n_grid_r"], g["n_grid_z"], g["r_size"], g["z_size"], 1e-12)
    sim.add_plasma(-1, 1.0, [0.004], [0.01], [3.5e8], [0.0], [0.0], [1.0])
    with pytest.raises(capi.Pdp3Error) as ei:
        sim.step(1)
    assert ei.value.code == capi.ERR_SUPERLUMINAL

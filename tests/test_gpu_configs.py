"""Reduced copies of the benchmark configurations against the FP64 oracle.

* cfg2 (electrons and mobile ions m_i = 1836, Vay, Gaussian 0.99c bunch,
  PML 16 cells) on a 32 x 256-cell grid at its 2.5e-4 m spacing, 20 ppc per
  species, a 10^4-macro bunch: the FP64 plain form within 1e-10 normwise of
  the oracle after 100 steps (fields and particle state), or 10x the run's
  self-change where one deposit-order-sensitive event moves it more, and the
  FP32 form's E_z and momentum errors against the same oracle run;
* cfg3 (the 0 -> 1e18 m^-3 density ramp, 100 ppc, Gaussian 0.999c bunch of
  sigma_r 0.5 mm, sigma_z 1 mm, peak 5e17 m^-3) on a 24 x 256-cell slab at its
  spacing with the ramp over the first 128 z-cells: the FP32 form's E_z
  relative RMS error against the oracle after 100 steps <= 1e-4, the
  precision target of the configuration;
* cfg4: each member of a three-instance ensemble of uneven grids (one graph
  replay per step, one particle launch) against the oracle run of
  that member alone, FP32 E_z relative RMS <= 1e-4 after 60 steps.

The FP32 runs start from the FP64 initial state rounded to the FP32 form
(cell index + FP32 offsets, FP32 u), the oracle from the FP64 state itself.
"""
import numpy as np
import pytest

from helpers import compare_fields, compare_species, normwise

pytestmark = pytest.mark.gpu

FP32_EZ_RMS_BAR = 1e-4


def rel_rms(a, ref):
    den = np.sqrt(np.mean(ref ** 2))
    return float(np.sqrt(np.mean((a - ref) ** 2)) / den)


def momentum_errors(st32, st64):
    """relative RMS and normwise max error of u = gamma v over the particles
    alive in both states (u from the downloaded velocities)"""
    m = st32["alive"].astype(bool) & st64["alive"].astype(bool)

    def u(st):
        v = np.stack([st["vr"][m], st["vphi"][m], st["vz"][m]])
        g = 1.0 / np.sqrt(1.0 - np.sum(v * v, axis=0) / 9e16)
        return v * g

    a, b = u(st32), u(st64)
    d = np.sqrt(np.sum((a - b) ** 2, axis=0))
    n = np.sqrt(np.sum(b * b, axis=0))
    return {"momentum_rel_rms": float(np.sqrt(np.mean(d ** 2)) / np.sqrt(np.mean(n ** 2))),
            "momentum_rel_max": float(np.max(d) / np.max(n)),
            "momentum_per_particle_p9999": float(np.quantile(d / np.maximum(n, 1e-300), 0.9999))}


def oracle_of(orc, w, plasma_states, threads=8):
    sim = orc.Sim(w.n_grid_r, w.n_grid_z, w.r_size, w.z_size, w.dt, pusher=w.pusher, threads=threads)
    pml = w.pml()
    if pml:
        sim.set_pml(*pml)
    for p, st in zip(w.plasma, plasma_states):
        sim.add_plasma(p.charge, p.mass, st["r"], st["z"], st["vr"], st["vphi"], st["vz"], st["w"], st["alive"])
    for b in w.bunches:
        sim.add_bunch(b.charge, b.mass, b.r, b.w, b.inject_u, b.duration, b.hole_duration, b.velocity, b.bunch_number)
    return sim


def cfg2_reduced():
    from pdp3_amd import workloads as wl

    w = wl.config("cfg2", cells_r=32, cells_z=256, bunch_macro=10000)
    for p in w.plasma:
        p.ppc = 20
    return w


def cfg3_reduced():
    from pdp3_amd import workloads as wl

    return wl.config("cfg3", cells_r=24, cells_z=256, bunch_macro=10000, ramp_cells=128)


@pytest.fixture(scope="module")
def cfg2_run(orc):
    """the reduced cfg2 on the GPU (FP64 form) and in the oracle, 100 steps"""
    from pdp3_amd import capi
    from pdp3_amd import workloads as wl

    w = cfg2_reduced()
    print("cfg2 reduced:", w.cells_r, "x", w.cells_z, "cells,", w.particle_count(), "particles, Vay, 100 steps",
          flush=True)
    gpu = wl.build_simulation(w, precision=capi.PREC_FP64)
    states = [gpu.download(j) for j in range(len(w.plasma))]
    ora = oracle_of(orc, w, states)
    gpu.step(100)
    ora.step(100)
    yield w, gpu, ora
    gpu.close()


def test_cfg2_reduced_fp64_parity(orc, cfg2_run):
    """Within 1e-10 normwise, or 10x the self-change of the run where the
    loop is that sensitive (capped at 2e-8 for fields, 5e-8 for particle
    state).  The self-change is the largest difference from the first runs
    (the oracle's and the GPU's) of the oracle with its serial deposit sums in
    other orders (reversed, shuffled particles) and of two more GPU runs
    (whose FP64 atomics sum in another order each time).  Most runs agree
    with the oracle to ~1e-13; about one GPU run in three moved J_r by
    3.1e-10 (a particle whose branch flips with the last bit of a deposit
    sum), which the other runs then differ from by the same amount."""
    from pdp3_amd import capi
    from pdp3_amd import workloads as wl

    w, gpu, ora = cfg2_run
    assert gpu.time == ora.time
    errs = compare_fields(gpu, ora)
    nsp = len(w.plasma) + len(w.bunches)
    sp = [compare_species(gpu.download(j), ora.get_species(j)) for j in range(nsp)]
    floor = {k: 0.0 for k in errs}
    sp_floor = [dict(r=0.0, z=0.0, vr=0.0, vphi=0.0, vz=0.0) for _ in range(nsp)]

    def widen(sim_fields, sim_species, ref_fields, ref_species):
        for k, e in compare_fields(sim_fields, ref_fields).items():
            floor[k] = max(floor[k], e)
        for j in range(nsp):
            for k, e in compare_species(sim_species(j), ref_species(j)).items():
                if k in sp_floor[j]:
                    sp_floor[j][k] = max(sp_floor[j][k], e)

    rng = np.random.default_rng(7)
    init = wl.build_simulation(w, precision=capi.PREC_FP64)
    states = [init.download(j) for j in range(len(w.plasma))]
    init.close()
    for variant in ("reverse", "shuffle"):
        perms = [rng.permutation(len(st["r"])) if variant == "shuffle" else np.arange(len(st["r"]))
                 for st in states]
        o2 = oracle_of(orc, w, [{k: v[p] for k, v in st.items()} for st, p in zip(states, perms)])
        o2.set_reverse_deposit(variant == "reverse")
        o2.step(100)

        def species_of(j, o2=o2, perms=perms):
            s2 = o2.get_species(j)
            if j < len(w.plasma):
                inv = np.argsort(perms[j])
                s2 = {k: v[inv] for k, v in s2.items()}
            return s2

        widen(o2, species_of, ora, ora.get_species)
    for _ in range(2):
        g2 = wl.build_simulation(w, precision=capi.PREC_FP64)
        g2.step(100)
        widen(g2, g2.download, gpu, gpu.download)
        g2.close()
    print("fields", errs)
    print("self-change floor", floor)
    for name, e in errs.items():
        assert e <= min(2e-8, max(1e-10, 10 * floor[name])), (name, e, floor[name])
    for j in range(nsp):
        print("species", j, sp[j], "floor", sp_floor[j])
        assert sp[j]["alive_mismatch"] == 0
        for k in ("r", "z", "vr", "vphi", "vz"):
            assert sp[j][k] <= min(5e-8, max(1e-10, 10 * sp_floor[j][k])), (j, k, sp[j][k], sp_floor[j][k])


def test_cfg2_reduced_fp32_error(cfg2_run):
    from pdp3_amd import capi
    from pdp3_amd import workloads as wl

    w, _, ora = cfg2_run
    g32 = wl.build_simulation(w, precision=capi.PREC_FP32)
    g32.step(100)
    ez = g32.field("E_z")
    ref = ora.field("E_z")
    out = {"E_z_rel_rms": rel_rms(ez, ref), "E_z_rel_max": normwise(ez, ref)}
    for j in range(len(w.plasma)):
        out[f"species{j}"] = momentum_errors(g32.download(j), ora.get_species(j))
    print("cfg2 reduced FP32 vs oracle after 100 steps", out)
    g32.close()
    assert out["E_z_rel_rms"] <= FP32_EZ_RMS_BAR, out
    for j in range(len(w.plasma)):
        assert out[f"species{j}"]["momentum_rel_rms"] <= 1e-4, out


def test_cfg3_reduced_fp32_error(orc):
    from pdp3_amd import capi
    from pdp3_amd import workloads as wl

    w = cfg3_reduced()
    print("cfg3 reduced:", w.cells_r, "x", w.cells_z, "cells,", w.particle_count(), "particles, 100 steps",
          flush=True)
    g64 = wl.build_simulation(w, precision=capi.PREC_FP64)
    states = [g64.download(0)]
    g64.close()
    ora = oracle_of(orc, w, states)
    g32 = wl.build_simulation(w, precision=capi.PREC_FP32, sort_interval=20)
    ora.step(100)
    g32.step(100)
    ez = g32.field("E_z")
    ref = ora.field("E_z")
    out = {"E_z_rel_rms": rel_rms(ez, ref), "E_z_rel_max": normwise(ez, ref)}
    # the FP32 state is cell sorted: match particles by their initial order
    # through a sort-free FP32 run for the momenta
    g32u = wl.build_simulation(w, precision=capi.PREC_FP32)
    g32u.step(100)
    out["E_z_rel_rms_unsorted"] = rel_rms(g32u.field("E_z"), ref)
    out["plasma"] = momentum_errors(g32u.download(0), ora.get_species(0))
    print("cfg3 reduced FP32 vs oracle after 100 steps", out)
    g32.close()
    g32u.close()
    assert out["E_z_rel_rms"] <= FP32_EZ_RMS_BAR, out
    assert out["E_z_rel_rms_unsorted"] <= FP32_EZ_RMS_BAR, out
    assert out["plasma"]["momentum_rel_rms"] <= 1e-4, out


def test_cfg4_members_against_oracle(orc):
    from pdp3_amd import capi
    from pdp3_amd import workloads as wl

    insts = [wl.Instance(cells_r=16, cells_z=128, ppc=12, density=1e17, bunch_ratio=0.2, seed=31, bunch_macro=3000),
             wl.Instance(cells_r=32, cells_z=64, ppc=5, density=3e16, bunch_ratio=0.5, seed=32, bunch_macro=2000),
             wl.Instance(cells_r=8, cells_z=256, ppc=20, density=1e18, bunch_ratio=0.05, seed=33, bunch_macro=4000)]
    print("cfg4 members vs oracle, 60 steps", flush=True)
    members = [wl.build_instance(i) for i in insts]
    ens = capi.Ensemble(members)
    ens.set_sort_interval(20)
    ens.step(60)
    ens.close()
    for inst, m in zip(insts, members):
        s64 = wl.build_instance(inst, precision=capi.PREC_FP64)
        st = s64.download(0)
        s64.close()
        ww = wl.Workload("cfg4", inst.cells_r, inst.cells_z, inst.dr, inst.dr, 60, inst.pml_cells, 0)
        ora = orc.Sim(ww.n_grid_r, ww.n_grid_z, ww.r_size, ww.z_size, ww.dt, pusher=0, threads=8)
        ora.set_pml(*ww.pml())
        ora.add_plasma(-1, 1.0, st["r"], st["z"], st["vr"], st["vphi"], st["vz"], st["w"], st["alive"])
        rng = np.random.default_rng(inst.seed + 1)
        ora.add_bunch_loaded(-1, 1.0, *wl.loaded_gaussian_bunch(rng, inst))
        ora.step(60)
        assert abs(m.time - ora.time) <= 1e-25
        errs = {n: rel_rms(m.field(n), ora.field(n)) for n in ("E_r", "E_z", "H_phi")}
        errs["E_z_rel_max"] = normwise(m.field("E_z"), ora.field("E_z"))
        print("member", inst.cells_r, "x", inst.cells_z, errs)
        assert errs["E_z"] <= FP32_EZ_RMS_BAR, errs

"""Ensembles: independent instances of different shapes advanced together
(graph replay, one particle launch) give the same result as each
instance stepped alone, with and without periodic cell sorts, for every tile
shape the autotuner may pick, with bunches loaded at the start or injected
over time (Bunch::bunch_inject from the device step counter)."""
import numpy as np
import pytest

from helpers import normwise

pytestmark = pytest.mark.gpu


def _instances():
    from pdp3_amd import workloads as wl

    return [wl.Instance(cells_r=16, cells_z=128, ppc=12, density=1e17, bunch_ratio=0.2, seed=11, bunch_macro=3000),
            wl.Instance(cells_r=32, cells_z=64, ppc=5, density=3e16, bunch_ratio=0.5, seed=12, bunch_macro=2000),
            wl.Instance(cells_r=8, cells_z=256, ppc=20, density=1e18, bunch_ratio=0.05, seed=13, bunch_macro=4000)]


def _compare(a, b, tol):
    for name in ("E_r", "E_z", "H_phi", "J_z", "rho_bunch"):
        e = normwise(a.field(name), b.field(name))
        assert e <= tol, (name, e)
    for j in range(2):
        sa, sb = a.download(j), b.download(j)
        # the cell sort may order particles of one cell differently in the two
        # runs: compare the states in the order of (r, z)
        oa, ob = np.lexsort((sa["z"], sa["r"])), np.lexsort((sb["z"], sb["r"]))
        assert np.array_equal(sa["alive"][oa], sb["alive"][ob])
        for k in ("r", "z", "vr", "vphi", "vz"):
            assert normwise(sa[k][oa], sb[k][ob]) <= tol, (j, k)


@pytest.mark.parametrize("sort_interval", [0, 5])
def test_ensemble_matches_single_instances(sort_interval):
    from pdp3_amd import capi
    from pdp3_amd import workloads as wl

    insts = _instances()
    print("ensemble of", len(insts), "sort_interval", sort_interval, flush=True)
    alone = [wl.build_instance(i) for i in insts]
    members = [wl.build_instance(i) for i in insts]
    for s in alone:
        s.set_sort_interval(sort_interval)
        s.step(12)
    ens = capi.Ensemble(members)
    ens.set_sort_interval(sort_interval)
    ens.step(12)
    ms, n = ens.timing()
    assert n == 12 and ms > 0
    for a, m in zip(alone, members):
        assert abs(a.time - m.time) == 0.0
        _compare(m, a, 1e-10)
    ens.close()


def test_ensemble_tiles_and_autotune():
    from pdp3_amd import capi
    from pdp3_amd import workloads as wl

    insts = _instances()
    ref = None
    for ppt, waves in [(16, 4), (16, 1), (8, 2), (4, 4)]:
        print("ensemble tile", ppt, waves, flush=True)
        members = [wl.build_instance(i) for i in insts]
        ens = capi.Ensemble(members)
        ens.set_tile(ppt, waves)
        assert ens.tile() == (ppt, waves)
        ens.step(8)
        ens.close()
        if ref is None:
            ref = members
            continue
        for a, b in zip(members, ref):
            _compare(a, b, 1e-10)
    members = [wl.build_instance(i) for i in insts]
    ens = capi.Ensemble(members)
    best = ens.autotune(1)
    assert best in [(16, 4), (16, 2), (16, 1), (8, 4), (8, 2), (4, 4)]
    ens.step(2)
    assert members[0].time > 0
    ens.close()


def test_ensemble_sort_orders_by_cell_or_band():
    """The ensemble's batched sort orders each instance's plasma stably by its
    key: row-major cells for instances with at least 32 particles per cell,
    bands of two rows column by column below that.  Cold particles with no
    fields stay where they are, so two steps (one sort each) leave the
    shuffled initial state in exactly the sorted order."""
    from pdp3_amd import capi

    rng = np.random.default_rng(41)
    dr = 5e-4
    shapes = [(16, 64, 6), (9, 40, 40), (33, 20, 3)]  # cells r, cells z, particles per cell
    sims, before, keys = [], [], []
    for cr, cz, ppc in shapes:
        sim = capi.Simulation(cr + 1, cz + 1, cr * dr, cz * dr, 0.5 * dr / 3e8, precision=capi.PREC_FP32)
        n = cr * cz * ppc
        # inside the reflecting walls at dr/2 from the axis and the boundaries
        r = rng.uniform(0.51, cr - 0.51, n) * dr
        z = rng.uniform(0.51, cz - 0.51, n) * dr
        r[: n // 8] = (rng.integers(1, cr - 1, n // 8) + 0.25) * dr  # many particles sharing cells
        zero = np.zeros(n)
        sim.add_plasma(-1, 1.0, r, z, zero, zero, zero, np.full(n, 1e3))
        st = sim.download(0)
        # the cells the library packed the particles into (CELL_NUMBER of the input)
        i = np.clip(np.ceil(r / sim.dr).astype(np.int64) - 1, 0, cr - 1)
        k = np.clip(np.ceil(z / sim.dz).astype(np.int64) - 1, 0, cz - 1)
        band = n < 32 * cr * cz
        keys.append((i // 2) * 2 * cz + 2 * k + (i & 1) if band else i * cz + k)
        before.append(st)
        sims.append(sim)
    ens = capi.Ensemble(sims)
    ens.set_sort_interval(1)
    ens.step(2)
    ens.close()
    for sim, st, key in zip(sims, before, keys):
        after = sim.download(0)
        order = np.argsort(key, kind="stable")
        for name in ("r", "z", "vr", "vphi", "vz", "w"):
            assert np.array_equal(after[name], st[name][order]), name
        sim.close()


def test_ensemble_rejects_fp64():
    from pdp3_amd import capi
    from pdp3_amd import workloads as wl

    inst = _instances()[0]
    s64 = wl.build_instance(inst, precision=capi.PREC_FP64)
    with pytest.raises(capi.Pdp3Error):
        capi.Ensemble([s64])


@pytest.mark.parametrize("sort_interval", [0, 10])
def test_ensemble_injected_bunches(sort_interval):
    """Instances whose bunches are injected over time (the 64 x 256
    configuration's uniform 0.9c bunch, on two grids, one with a two-bunch
    train) beside a loaded-bunch instance: the ensemble's device-side injection
    follows each instance's own clock exactly as its single stepping does."""
    from pdp3_amd import capi
    from pdp3_amd import workloads as wl

    def injected(cells_r, cells_z, seed, train):
        w = wl.config("cfg1", seed=seed, cells_r=cells_r, cells_z=cells_z, pml_cells=4, bunch_macro=3000)
        if train:
            b0 = w.bunches[0]
            b0.hole_duration = 0.4 * b0.duration
            b1 = wl.Bunch(b0.charge, b0.mass, b0.r.copy(), b0.w.copy(), b0.inject_u[::-1].copy(), b0.duration,
                          b0.hole_duration, b0.velocity, 1)
            w.bunches.append(b1)
        return w

    ws = [injected(16, 96, 41, False), injected(24, 64, 42, True)]
    print("ensemble with injected bunches, sort_interval", sort_interval, flush=True)
    alone = [wl.build_simulation(w, precision=capi.PREC_FP32) for w in ws]
    members = [wl.build_simulation(w, precision=capi.PREC_FP32) for w in ws]
    inst = _instances()[0]
    alone.append(wl.build_instance(inst))
    members.append(wl.build_instance(inst))
    for s in alone:
        s.set_sort_interval(sort_interval)
        s.step(60)
    ens = capi.Ensemble(members)
    ens.set_sort_interval(sort_interval)
    ens.step(60)
    ens.close()
    for a, m in zip(alone, members):
        assert abs(a.time - m.time) == 0.0
        assert np.max(np.abs(m.field("rho_bunch"))) > 0
        _compare(m, a, 1e-10)
        for j in range(1, a.n_species()):  # the bunches (plain layout: same order in both)
            sa, sm = a.download(j), m.download(j)
            assert np.array_equal(sa["alive"], sm["alive"]), j
            assert sa["alive"].sum() > 0
            for k in ("r", "z", "vr", "vphi", "vz"):
                assert normwise(sm[k], sa[k]) <= 1e-10, (j, k)

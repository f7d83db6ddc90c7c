"""The cell sort: the resulting order of the living particles equals a stable
sort by the reference's cell index (CELL_NUMBER clamped at 0,
include/geometry.h:44) of the state before the sort; dead particles move to
the end; no particle is lost or changed; sorting a sorted state changes
nothing; a step after a sort gives the same fields as without it (up to the
order of the grid sums)."""
import numpy as np
import pytest

from helpers import normwise

pytestmark = pytest.mark.gpu


def _cell_keys(st, dr, dz, nr, nz):
    i = np.clip(np.ceil(st["r"] / dr).astype(np.int64) - 1, 0, nr - 2)
    k = np.clip(np.ceil(st["z"] / dz).astype(np.int64) - 1, 0, nz - 2)
    return i * (nz - 1) + k


@pytest.mark.parametrize("precision", [0, 1])
def test_sort_is_stable_by_cell(precision):
    from pdp3_amd import capi

    rng = np.random.default_rng(31)
    nr, nz, dr = 65, 257, 5e-4
    sim = capi.Simulation(nr, nz, 64 * dr, 256 * dr, 0.5 * dr / 3e8, precision=precision)
    n = 300000
    r = rng.uniform(dr / 2, 63.5 * dr, n)
    z = rng.uniform(dr / 2, 255.5 * dr, n)
    # many exact ties and particles on cell faces
    r[:50000] = (rng.integers(1, 60, 50000) + 0.25) * dr
    z[50000:60000] = rng.integers(1, 250, 10000) * dr
    v = rng.standard_normal((3, n)) * 1e5
    w = rng.uniform(1, 2, n)
    alive = (rng.random(n) > 0.05).astype(np.uint8)
    idx = sim.add_plasma(-1, 1.0, r, z, v[0], v[1], v[2], w, alive)
    before = sim.download(idx)
    key = _cell_keys(before, sim.dr, sim.dz, nr, nz)
    key = np.where(before["alive"] == 1, key, (nr - 1) * (nz - 1))
    order = np.argsort(key, kind="stable")
    sim.sort()
    after = sim.download(idx)
    for k in ("r", "z", "vr", "vphi", "vz", "w", "alive"):
        assert np.array_equal(after[k], before[k][order]), k
    sim.sort()
    again = sim.download(idx)
    for k in ("r", "z", "vr", "vphi", "vz", "w", "alive"):
        assert np.array_equal(again[k], after[k]), k


def test_sorted_step_matches_unsorted():
    from pdp3_amd import capi
    from pdp3_amd import workloads as wl

    w = wl.config("cfg1")
    w.bunches = []
    a = wl.build_simulation(w)
    b = wl.build_simulation(w)
    a.step(20)
    b.step(20)
    b.sort()
    a.step(5)
    b.step(5)
    for name in ("E_r", "E_z", "H_phi", "J_r", "J_z", "J_phi"):
        e = normwise(b.field(name), a.field(name))
        assert e <= 1e-9, (name, e)


@pytest.mark.parametrize("precision", [0, 1])
def test_periodic_sort_of_unsorted_state(precision):
    """Periodic sorts in the loop size their record space from the previous
    sort: on a shuffled state the first finds more runs than that space and is
    skipped (the state keeps its order), the next ones sort.  The fields and
    the particles match the unsorted run; an explicit sort afterwards orders
    the state by cell."""
    from pdp3_amd import capi

    rng = np.random.default_rng(33)
    nr, nz, dr = 33, 129, 5e-4
    sims = []
    for si in (0, 2):
        sim = capi.Simulation(nr, nz, 32 * dr, 128 * dr, 0.5 * dr / 3e8, precision=precision)
        sims.append(sim)
    n = 200000
    r = rng.uniform(dr / 2, 31.5 * dr, n)
    z = rng.uniform(dr / 2, 127.5 * dr, n)
    v = rng.standard_normal((3, n)) * 1e5
    w = rng.uniform(1, 2, n)
    for sim, si in zip(sims, (0, 2)):
        sim.boundary_conditions()
        sim.add_plasma(-1, 1.0, r, z, v[0], v[1], v[2], w)
        sim.set_sort_interval(si)
        sim.step(7)  # sorts before steps 3, 5 and 7
    for name in ("E_r", "E_z", "H_phi", "J_r", "J_z", "J_phi"):
        e = normwise(sims[1].field(name), sims[0].field(name))
        assert e <= 1e-9, (name, e)
    a, b = sims[0].download(0), sims[1].download(0)
    oa, ob = np.lexsort((a["z"], a["r"])), np.lexsort((b["z"], b["r"]))
    for k in ("r", "z", "vr", "vphi", "vz", "w"):
        assert normwise(b[k][ob], a[k][oa]) <= 1e-9, k
    sims[1].step(1)
    sims[1].set_sort_interval(0)
    sims[1].sort()
    st = sims[1].download(0)
    key = _cell_keys(st, sims[1].dr, sims[1].dz, nr, nz)
    assert np.all(np.diff(key) >= 0)

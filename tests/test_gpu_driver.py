"""The drop-in driver end to end: `pdp3 -f parameters.xml` on the reference's
functional smoke configuration (test/functional/parameters.xml.tmpl.in) on the
GPU, its probe files compared with the reference's own goldens
(test/functional/true_data/*/4.dat.gz).

FP64 plain form: every field frame to the goldens' 6 printed digits, and the
reference's own acceptance (np.allclose, rtol 1e-3, atol 1, NaNs removed,
test/functional/test.py:152-171) for all twelve probes including the two
temperature frames.  FP32 form (through the in-process entry point): the same
frames to 1e-3 normwise.  The temperature and density probes run on the
device (pdp3_temperature, pdp3_density); the density frame is compared with
the oracle's ParticlesDensity on the same state.
"""
import os
import subprocess

import numpy as np
import pytest

import smoke_case as sc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TEMPLATE = os.path.join(HERE, "golden", "smoke", "parameters.xml.in")
GOLDEN = os.path.join(HERE, "golden", "smoke")


# a density probe (src/loadInitParam.cpp:381-390) beside the template's own probes
DENSITY_PROBE = ('<probe type="frame" r_start="100" z_start="100" r_end="150" z_end="150" schedule="5" '
                 'component="density" specie="Electrons" />')


def _params(tmp_path, precision):
    text = open(TEMPLATE).read().replace("@RESULT_PATH@", str(tmp_path)).replace("@PRECISION@", precision)
    text = text.replace("</probes>", DENSITY_PROBE + "\n    </probes>")
    p = tmp_path / "parameters.xml"
    p.write_text(text)
    return str(p)


def _read(path):
    with open(path) as f:
        return np.array(f.read().split(), dtype=np.float64)


def _frame_path(root, comp, specie=None):
    r0, z0, r1, z1 = sc.FRAMES.get(comp, (100, 100, 150, 150))
    sub = comp if specie is None else f"{comp}/{specie}"
    return os.path.join(root, "pdp3_result", sub, f"frame_{r0}:{z0}_{r1}:{z1}", f"{sc.FRAME_NO}.dat")


def _golden_t(specie):
    import gzip
    with gzip.open(os.path.join(GOLDEN, f"T__{specie}__frame_100_100_150_150__4.dat.gz"), "rb") as f:
        return np.array(f.read().split(), dtype=np.float64)


def _reference_accepts(ours, ref):
    """test/functional/test.py:152-171: NaNs dropped, np.allclose(rtol 1e-3, atol 1)."""
    ours = ours[~np.isnan(ours)]
    ref = ref[~np.isnan(ref)]
    return ours.shape == ref.shape and np.allclose(ours, ref, rtol=1e-3, atol=1)


@pytest.fixture(scope="module")
def fp64_run(tmp_path_factory):
    from pdp3_amd import driver

    tmp = tmp_path_factory.mktemp("smoke64")
    print("pdp3 driver, smoke configuration, fp64", flush=True)
    r = subprocess.run([driver.DRIVER, "-f", _params(tmp, "fp64")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return tmp, r.stdout


@pytest.mark.parametrize("comp", list(sc.FRAMES))
def test_driver_fp64_frames(fp64_run, comp):
    root, _ = fp64_run
    ours = _read(_frame_path(root, comp))
    ref = sc.golden(comp)
    assert ours.size == ref.size
    ours = ours.reshape(ref.shape)
    rel_max, small_max, _ = sc.compare(ours, ref)
    # the files hold 6 significant digits, as the goldens do
    assert rel_max < 2e-5 and small_max < 1e-9, (comp, rel_max, small_max)
    assert _reference_accepts(ours.ravel(), ref.ravel())


@pytest.mark.parametrize("specie", ["Electrons", "Ions"])
def test_driver_fp64_temperature(fp64_run, specie):
    root, _ = fp64_run
    ours = _read(_frame_path(root, "T", specie))
    ref = _golden_t(specie)
    assert ours.size == ref.size == 50 * 50
    # empty cells (0/0) are NaN in both
    assert np.array_equal(np.isnan(ours), np.isnan(ref))
    assert _reference_accepts(ours, ref)
    # cells with two or more particles carry a physical temperature
    big = np.abs(ref) > 1e-6
    assert big.sum() > 10
    assert np.max(np.abs(ours[big] - ref[big]) / np.abs(ref[big])) < 1e-3


def test_driver_fp64_density(fp64_run, orc):
    """The density probe frame against ParticlesDensity::calc_density of the
    oracle on the same loop state (the smoke replay to the fifth dump)."""
    root, _ = fp64_run
    ours = _read(_frame_path(root, "density", "Electrons")).reshape(50, 50)
    plasma, bunch = sc.initial_state(orc)
    g, T = sc.GEOM, sc.TIME
    sim = orc.Sim(g["n_grid_r"], g["n_grid_z"], g["r_size"], g["z_size"], T["dt"],
                  current_time=T["current_time"], end_time=T["end_time"])
    for p in plasma:
        sim.add_plasma(p["charge"], p["mass"], p["r"], p["z"], p["vr"], p["vphi"], p["vz"], p["w"])
    b = bunch
    sim.add_bunch(b["charge"], b["mass"], b["r"], b["w"], b["inject_u"], b["duration"], b["hole_duration"],
                  b["velocity"], b["bunch_number"])
    dumps, ref = 0, None
    while ref is None:
        fires = sc.probe_fires(sim.time, T["dt"])
        sim.step(1)
        if fires:
            if dumps == sc.FRAME_NO:
                ref = sc.t_frame(sim.density(0))
            dumps += 1
    assert np.count_nonzero(ref) > 100
    # the file holds 6 significant digits
    big = np.abs(ref) > 1e-6 * np.max(np.abs(ref))
    assert np.max(np.abs(ours[big] - ref[big]) / np.abs(ref[big])) < 2e-5
    assert np.max(np.abs(ours[~big])) <= 1e-6 * np.max(np.abs(ref))


def test_driver_fp64_layout(fp64_run):
    root, out = fp64_run
    base = os.path.join(root, "pdp3_result")
    # 23 loop iterations, probes fire at t/dt = 0, 5, 10, 15, 20: files 0..4
    for comp, sub in (("E_z", "col_100"), ("E_r", "row_10"), ("rho_beam", "dot_10_50")):
        files = sorted(os.listdir(os.path.join(base, comp, sub)))
        assert files == [f"{i}.dat" for i in range(5)], (comp, files)
    col = _read(os.path.join(base, "E_z", "col_100", "4.dat"))
    row = _read(os.path.join(base, "E_r", "row_10", "4.dat"))
    dot = _read(os.path.join(base, "rho_beam", "dot_10_50", "4.dat"))
    assert col.size == 255 and row.size == 2047 and dot.size == 1
    ez = _read(_frame_path(root, "E_z")).reshape(32, 256)        # rows 64..95, cols 512..767
    er = _read(_frame_path(root, "E_r")).reshape(32, 256)        # rows 0..31, cols 0..255
    rho = _read(_frame_path(root, "rho_beam")).reshape(254, 2046)
    assert np.array_equal(row[:256], er[10])
    assert dot[0] == rho[10, 50]
    assert np.all(np.isfinite(col)) and col.shape == (255,)
    assert ez.shape == (32, 256)
    assert "Simulation Completed" in out


def test_driver_fp32(tmp_path):
    from pdp3_amd import driver

    print("pdp3 driver in-process, smoke configuration, fp32", flush=True)
    steps = driver.run(_params(tmp_path, "fp32"))
    assert steps == 23
    for comp in ("E_r", "E_z", "H_phi", "J_z", "rho_beam"):
        ref = sc.golden(comp).ravel()
        ours = _read(_frame_path(tmp_path, comp))
        e = np.linalg.norm(ours - ref) / np.linalg.norm(ref)
        print(comp, "fp32 vs golden normwise", e)
        assert e < 1e-3, (comp, e)

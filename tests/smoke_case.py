"""The reference's functional smoke case, rebuilt from its template.

test/functional/parameters.xml.tmpl.in:8-68 (geometry, pml, time, species,
beam) and the probes at :88-103.  The golden frames in tests/golden/smoke are
the reference's test/functional/true_data/*/frame_*/4.dat.gz, i.e. the fifth
probe dump (test/functional/test.py:479-487).
"""
import gzip
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "smoke")

EL_CHARGE = 1.6e-19
EL_MASS = 9.1e-31

GEOM = dict(r_size=0.25, z_size=2.0, n_grid_r=255, n_grid_z=2047)
TIME = dict(current_time=0.0, end_time=2.2e-11, dt=1e-12)
SPECIES = [
    dict(name="Electrons", charge=-1, mass=1, number=100000, n1=2e16, n2=2.05e16, temperature=1.0),
    dict(name="Ions", charge=1, mass=1836, number=100000, n1=2e16, n2=2.05e16, temperature=0.1),
]
BEAM = dict(charge=-1, mass=1, velocity=2e8, number_bunches=1, bunches_distance=0.02,
            number=1000000, length=2.0, radius=0.02, density=1e15)
SCHEDULE = 5
FRAME_NO = 4
# component -> (r_start, z_start, r_end, z_end)
FRAMES = {
    "E_r": (0, 0, 32, 256), "E_phi": (32, 256, 64, 512), "E_z": (64, 512, 96, 768),
    "H_r": (96, 768, 128, 1024), "H_phi": (128, 1024, 160, 1280), "H_z": (160, 1280, 192, 2046),
    "J_r": (192, 1536, 244, 1792), "J_phi": (244, 1792, 255, 2047), "J_z": (0, 0, 32, 256),
    "rho_beam": (0, 0, 254, 2046),
}


def golden(comp):
    r0, z0, r1, z1 = FRAMES[comp]
    fn = os.path.join(GOLDEN, f"{comp}__frame_{r0}_{z0}_{r1}_{z1}__{FRAME_NO}.dat.gz")
    with gzip.open(fn, "rb") as f:
        data = np.array(f.read().split(), dtype=np.float64)
    return data.reshape(r1 - r0, z1 - z0)


# temperature probes: (r_start, z_start, r_end, z_end) and species names
T_FRAME = (100, 100, 150, 150)


def golden_t(specie):
    r0, z0, r1, z1 = T_FRAME
    fn = os.path.join(GOLDEN, f"T__{specie}__frame_{r0}_{z0}_{r1}_{z1}__{FRAME_NO}.dat.gz")
    with gzip.open(fn, "rb") as f:
        data = np.array([float(x) for x in f.read().split()], dtype=np.float64)
    return data.reshape(r1 - r0, z1 - z0)


def t_frame(arr):
    r0, z0, r1, z1 = T_FRAME
    return arr[r0:r1, z0:z1]


def dr_dz():
    return (GEOM["r_size"] / (GEOM["n_grid_r"] - 1), GEOM["z_size"] / (GEOM["n_grid_z"] - 1))


def initial_state(orc):
    """Initial particles exactly as the reference builds them in TESTMODE,
    produced with the oracle's loaders (src/particles.cpp:228-331,
    src/bunch.cpp:6-80)."""
    dr, dz = dr_dz()
    plasma = []
    for sp in SPECIES:
        r, z, w = orc.load_cylindrical(sp["number"], GEOM["r_size"], GEOM["z_size"], dr, dz, sp["n1"], sp["n2"], 0.0)
        vr, vphi, vz = orc.velocity_testmode(sp["number"], sp["mass"] * EL_MASS, sp["temperature"])
        plasma.append(dict(charge=sp["charge"], mass=sp["mass"], r=r, z=z, vr=vr, vphi=vphi, vz=vz, w=w))
    b = BEAM
    duration = b["length"] / b["velocity"]
    hole = b["bunches_distance"] / b["velocity"]
    r, w = orc.bunch_testmode(b["number"], b["radius"], dr, dz, duration, b["velocity"], b["density"])
    u = np.array([orc.random_reverse(i, 11) for i in range(b["number"])])
    bunch = dict(charge=b["charge"], mass=b["mass"], r=r, w=w, inject_u=u, duration=duration,
                 hole_duration=hole, velocity=b["velocity"], bunch_number=0)
    return plasma, bunch


def probe_fires(t, dt):
    """src/loadInitParam.cpp:329: (int)(current_time / delta_t) % schedule == 0."""
    return int(t / dt) % SCHEDULE == 0


def frame(arr, comp):
    r0, z0, r1, z1 = FRAMES[comp]
    return arr[r0:r1, z0:z1]


def compare(ours, ref, rtol=1e-3):
    """Max relative error over entries whose golden magnitude is above 1e-6 of
    the frame's largest, plus the max absolute error on the rest scaled by that
    largest magnitude."""
    scale = np.max(np.abs(ref)) if ref.size else 0.0
    big = np.abs(ref) > 1e-6 * scale
    rel = np.abs(ours[big] - ref[big]) / np.abs(ref[big]) if big.any() else np.zeros(1)
    small = np.abs(ours[~big] - ref[~big]) / scale if (~big).any() and scale > 0 else np.zeros(1)
    return float(rel.max()), float(small.max()), float(np.quantile(rel, 0.999))

"""ctypes wrapper around the CPU oracle (liboracle_pdp3.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product package never
imports it.

The oracle is a clean-room FP64 restatement of the reference loop
(cosmonaut-ok/pdp3); see pdp3_oracle.c for per-function citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_pdp3.so")

E_R, E_PHI, E_Z, H_R, H_PHI, H_Z, HH_R, HH_PHI, HH_Z, J_R, J_PHI, J_Z, RHO_BUNCH, SIGMA, EPSILON = range(15)
COMPONENTS = {
    "E_r": E_R, "E_phi": E_PHI, "E_z": E_Z,
    "H_r": H_R, "H_phi": H_PHI, "H_z": H_Z,
    "HH_r": HH_R, "HH_phi": HH_PHI, "HH_z": HH_Z,
    "J_r": J_R, "J_phi": J_PHI, "J_z": J_Z,
    "rho_bunch": RHO_BUNCH, "sigma": SIGMA, "epsilon": EPSILON,
}
PUSHER_BORIS, PUSHER_VAY = 0, 1

_D = C.POINTER(C.c_double)
_U8 = C.POINTER(C.c_uint8)
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, IEEE flags)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_sim_new.restype = C.c_void_p
        L.or_sim_new.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
        L.or_sim_free.argtypes = [C.c_void_p]
        L.or_sim_set_time.argtypes = [C.c_void_p] + [C.c_double] * 5
        L.or_sim_set_pml.argtypes = [C.c_void_p] + [C.c_double] * 5
        L.or_sim_set_pusher.argtypes = [C.c_void_p, C.c_int]
        L.or_sim_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.or_sim_set_reverse_deposit.argtypes = [C.c_void_p, C.c_int]
        L.or_set_gamma_inv_nearest.argtypes = [C.c_int]
        L.or_sim_time.restype = C.c_double
        L.or_sim_time.argtypes = [C.c_void_p]
        L.or_sim_dr.restype = C.c_double
        L.or_sim_dr.argtypes = [C.c_void_p]
        L.or_sim_dz.restype = C.c_double
        L.or_sim_dz.argtypes = [C.c_void_p]
        L.or_sim_field.restype = _D
        L.or_sim_field.argtypes = [C.c_void_p, C.c_int]
        L.or_sim_add_plasma.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int] + [_D] * 6 + [_U8]
        L.or_sim_add_bunch_loaded.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, _D, _D, _D, _D, _D, _D]
        L.or_sim_add_bunch.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, _D, _D, _D,
                                       C.c_double, C.c_double, C.c_double, C.c_int]
        L.or_sim_num_species.argtypes = [C.c_void_p]
        L.or_sim_species_count.argtypes = [C.c_void_p, C.c_int]
        L.or_sim_get_species.argtypes = [C.c_void_p, C.c_int] + [_D] * 7 + [_U8]
        L.or_sim_set_species.argtypes = [C.c_void_p, C.c_int] + [_D] * 5 + [_U8]
        L.or_sim_step.argtypes = [C.c_void_p, C.c_int]
        for f in ("or_sim_h_calc", "or_sim_e_calc", "or_sim_e_boundary_conditions", "or_sim_reset_j"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.or_sim_step_v.argtypes = [C.c_void_p, C.c_int]
        L.or_sim_full_current_distribution.argtypes = [C.c_void_p, C.c_int]
        L.or_sim_charge_weighting.argtypes = [C.c_void_p, C.c_int, _D]
        L.or_temperature.argtypes = [C.c_void_p, C.c_int, _D]
        L.or_density.argtypes = [C.c_void_p, C.c_int, _D]
        L.or_sim_bunch_inject.argtypes = [C.c_void_p, C.c_int]
        L.or_sim_get_e.argtypes = [C.c_void_p, C.c_double, C.c_double, _D]
        L.or_sim_get_h.argtypes = [C.c_void_p, C.c_double, C.c_double, _D]
        L.or_boris_velocity.argtypes = [C.c_double, C.c_double, C.c_double, _D, _D, _D]
        L.or_vay_velocity.argtypes = [C.c_double, C.c_double, C.c_double, _D, _D, _D]
        L.or_current_deposit.argtypes = [C.c_void_p] + [C.c_double] * 5 + [_D, _D]
        L.or_azimuthal_deposit.argtypes = [C.c_void_p] + [C.c_double] * 4 + [_D]
        L.or_charge_deposit.argtypes = [C.c_void_p] + [C.c_double] * 3 + [_D]
        L.or_cell_number.argtypes = [C.c_double, C.c_double]
        L.or_random_reverse.restype = C.c_double
        L.or_random_reverse.argtypes = [C.c_double, C.c_int]
        L.or_load_cylindrical.argtypes = [C.c_int] + [C.c_double] * 7 + [_D, _D, _D]
        L.or_velocity_distribution_testmode.argtypes = [C.c_int, C.c_double, C.c_double, _D, _D, _D]
        L.or_bunch_ctor_testmode.argtypes = [C.c_int] + [C.c_double] * 6 + [_D, _D]
        L.or_mj_table.restype = _D
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Sim:
    """Geometry + fields + species list + time, advanced by the oracle."""

    def __init__(self, n_grid_r, n_grid_z, r_size, z_size, dt, pusher=PUSHER_BORIS,
                 current_time=0.0, end_time=1e300, threads=1):
        L = lib()
        self.nr, self.nz = int(n_grid_r), int(n_grid_z)
        self.h = L.or_sim_new(self.nr, self.nz, float(r_size), float(z_size))
        if not self.h:
            raise ValueError("bad geometry")
        L.or_sim_set_time(self.h, current_time, 0.0, 0.0, end_time, dt)
        L.or_sim_set_pusher(self.h, pusher)
        L.or_sim_set_threads(self.h, threads)
        self.dt = dt
        self.dr = L.or_sim_dr(self.h)
        self.dz = L.or_sim_dz(self.h)
        self.species = []

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_sim_free(self.h)
            self.h = None

    def set_pml(self, l1, l2, l3, s1, s2):
        lib().or_sim_set_pml(self.h, l1, l2, l3, s1, s2)

    def set_reverse_deposit(self, reverse=True):
        lib().or_sim_set_reverse_deposit(self.h, 1 if reverse else 0)

    @property
    def time(self):
        return lib().or_sim_time(self.h)

    def field(self, name):
        """A live (n_r, n_z) float64 view of one grid array."""
        comp = COMPONENTS[name] if isinstance(name, str) else int(name)
        p = lib().or_sim_field(self.h, comp)
        return np.ctypeslib.as_array(p, shape=(self.nr, self.nz))

    def add_plasma(self, charge_e, mass_e, r, z, vr, vphi, vz, w, alive=None):
        r, z, vr, vphi, vz, w = map(_f64, (r, z, vr, vphi, vz, w))
        a = None if alive is None else np.ascontiguousarray(alive, dtype=np.uint8)
        idx = lib().or_sim_add_plasma(self.h, int(charge_e), float(mass_e), len(r), _dp(r), _dp(z), _dp(vr),
                                      _dp(vphi), _dp(vz), _dp(w), None if a is None else a.ctypes.data_as(_U8))
        if idx < 0:
            raise RuntimeError("too many species")
        self.species.append(idx)
        return idx

    def add_bunch(self, charge_e, mass_e, r, w, inject_u, duration, hole_duration, velocity, bunch_number=0):
        r, w, u = map(_f64, (r, w, inject_u))
        idx = lib().or_sim_add_bunch(self.h, int(charge_e), float(mass_e), len(r), _dp(r), _dp(w), _dp(u),
                                     duration, hole_duration, velocity, int(bunch_number))
        if idx < 0:
            raise RuntimeError("too many species")
        self.species.append(idx)
        return idx

    def add_bunch_loaded(self, charge_e, mass_e, r, z, vr, vphi, vz, w):
        arrs = [_f64(a) for a in (r, z, vr, vphi, vz, w)]
        idx = lib().or_sim_add_bunch_loaded(self.h, int(charge_e), float(mass_e), len(arrs[0]),
                                            *(_dp(a) for a in arrs))
        if idx < 0:
            raise RuntimeError("too many species")
        self.species.append(idx)
        return idx

    def get_species(self, idx):
        n = lib().or_sim_species_count(self.h, idx)
        out = {k: np.empty(n) for k in ("r", "z", "vr", "vphi", "vz", "q", "m")}
        alive = np.empty(n, dtype=np.uint8)
        lib().or_sim_get_species(self.h, idx, *(_dp(out[k]) for k in ("r", "z", "vr", "vphi", "vz", "q", "m")),
                                 alive.ctypes.data_as(_U8))
        out["alive"] = alive
        return out

    def temperature(self, idx):
        """TemperatureClassic of species idx (nr x nz, eV)"""
        out = np.zeros((self.nr, self.nz))
        st = lib().or_temperature(self.h, idx, _dp(out))
        if st != 0:
            raise RuntimeError(f"oracle temperature failed with status {st}")
        return out

    def density(self, idx):
        """ParticlesDensity::calc_density of species idx (nr x nz)"""
        out = np.zeros((self.nr, self.nz))
        st = lib().or_density(self.h, idx, _dp(out))
        if st != 0:
            raise RuntimeError(f"oracle density failed with status {st}")
        return out

    def step(self, n=1):
        st = lib().or_sim_step(self.h, int(n))
        if st != 0:
            raise RuntimeError(f"oracle step failed with status {st}")

    def get_e(self, r, z):
        out = np.empty(3)
        lib().or_sim_get_e(self.h, float(r), float(z), _dp(out))
        return out

    def get_h(self, r, z):
        out = np.empty(3)
        lib().or_sim_get_h(self.h, float(r), float(z), _dp(out))
        return out


def random_reverse(v, power):
    return lib().or_random_reverse(float(v), int(power))


def load_cylindrical(n, r_size, z_size, dr, dz, n1, n2, left=0.0):
    r, z, w = np.empty(n), np.empty(n), np.empty(n)
    lib().or_load_cylindrical(n, r_size, z_size, dr, dz, n1, n2, left, _dp(r), _dp(z), _dp(w))
    return r, z, w


def velocity_testmode(n, mass_kg, temperature):
    vr, vphi, vz = np.empty(n), np.empty(n), np.empty(n)
    st = lib().or_velocity_distribution_testmode(n, mass_kg, temperature, _dp(vr), _dp(vphi), _dp(vz))
    if st != 0:
        raise ValueError("unsupported temperature branch")
    return vr, vphi, vz


def bunch_testmode(n, radius, dr, dz, duration, velocity, density):
    r, w = np.empty(n), np.empty(n)
    lib().or_bunch_ctor_testmode(n, radius, dr, dz, duration, velocity, density, _dp(r), _dp(w))
    return r, w


def mj_table():
    return np.ctypeslib.as_array(lib().or_mj_table(), shape=(1000,)).copy()


def boris_velocity(q, m, dt, e, h, v):
    v = _f64(v).copy()
    st = lib().or_boris_velocity(q, m, dt, _dp(_f64(e)), _dp(_f64(h)), _dp(v))
    if st:
        raise ValueError(f"status {st}")
    return v


def vay_velocity(q, m, dt, e, h, v):
    v = _f64(v).copy()
    st = lib().or_vay_velocity(q, m, dt, _dp(_f64(e)), _dp(_f64(h)), _dp(v))
    if st:
        raise ValueError(f"status {st}")
    return v

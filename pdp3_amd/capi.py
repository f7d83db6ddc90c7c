"""ctypes binding of the C ABI in include/pdp3_hip.h (libpdp3_hip.so).

The HIP library is the only compute path: if it is missing this module raises
instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpdp3_hip.so")

OK = 0
ERR_BAD_ARG, ERR_HIP, ERR_OOM, ERR_NONFINITE, ERR_SUPERLUMINAL, ERR_STATE = -1, -2, -3, -4, -5, -6
PREC_FP64, PREC_FP32 = 0, 1
PUSHER_BORIS, PUSHER_VAY = 0, 1
SPECIES_PLASMA, SPECIES_BUNCH = 0, 1
COMPONENTS = {
    "E_r": 0, "E_phi": 1, "E_z": 2, "H_r": 3, "H_phi": 4, "H_z": 5,
    "HH_r": 6, "HH_phi": 7, "HH_z": 8, "J_r": 9, "J_phi": 10, "J_z": 11,
    "rho_bunch": 12, "sigma": 13, "epsilon": 14,
}
KERNEL_PARTICLES, KERNEL_H, KERNEL_E, KERNEL_SORT, KERNEL_BUNCH = range(5)
NOALLOC_MAX_NODES = 4718592  # PDP3_NOALLOC_MAX_NODES
TANGENT_DIRS = 4  # PDP3_TANGENT_DIRS

_D = C.POINTER(C.c_double)
_U8 = C.POINTER(C.c_uint8)
_P = C.c_void_p
_I = C.c_int
_F = C.c_double

# name -> (restype, argtypes); also the list of symbols include/pdp3_hip.h declares
SIGNATURES = {
    "pdp3_ctx_create": (_I, [_I, _I, _F, _F, _F, _F, _F, _F, _F, _I, _I, C.POINTER(_P)]),
    "pdp3_ctx_destroy": (_I, [_P]),
    "pdp3_last_error": (C.c_char_p, [_P]),
    "pdp3_geometry_get": (_I, [_P, _D]),
    "pdp3_time_get": (_I, [_P, _D, _D]),
    "pdp3_time_set": (_I, [_P, _F]),
    "pdp3_geometry_set_pml": (_I, [_P, _F, _F, _F, _F, _F]),
    "pdp3_efield_boundary_conditions": (_I, [_P]),
    "pdp3_efield_set_homogeneous": (_I, [_P, _F, _F, _F]),
    "pdp3_hfield_set_homogeneous": (_I, [_P, _F, _F, _F]),
    "pdp3_boundary_specify_initial_field": (_I, [_P, _F, _F, _F]),
    "pdp3_hfield_calc": (_I, [_P]),
    "pdp3_efield_calc": (_I, [_P]),
    "pdp3_current_reset": (_I, [_P]),
    "pdp3_rho_bunch_reset": (_I, [_P]),
    "pdp3_efield_get_field": (_I, [_P, _D, _D, _I, _D]),
    "pdp3_hfield_get_field": (_I, [_P, _D, _D, _I, _D]),
    "pdp3_grid_download": (_I, [_P, _I, _D]),
    "pdp3_grid_upload": (_I, [_P, _I, _D]),
    "pdp3_species_add": (_I, [_P, _I, _F, _I, C.POINTER(_I)]),
    "pdp3_bunch_add": (_I, [_P, _I, _F, _I, _D, _D, _D, _F, _F, _F, _I, C.POINTER(_I)]),
    "pdp3_species_upload": (_I, [_P, _I, _D, _D, _D, _D, _D, _D, _U8]),
    "pdp3_species_download": (_I, [_P, _I, _D, _D, _D, _D, _D, _D, _U8]),
    "pdp3_species_count": (_I, [_P, _I, C.POINTER(_I)]),
    "pdp3_species_init_uniform": (_I, [_P, _I, C.c_uint64, _I, _F, _I, _F]),
    "pdp3_particles_step_v": (_I, [_P, _I]),
    "pdp3_particles_full_current_distribution": (_I, [_P, _I]),
    "pdp3_particles_charge_weighting": (_I, [_P, _I]),
    "pdp3_bunch_inject": (_I, [_P, _I]),
    "pdp3_step": (_I, [_P, _I]),
    "pdp3_sync": (_I, [_P]),
    "pdp3_sort_particles": (_I, [_P]),
    "pdp3_set_sort_interval": (_I, [_P, _I]),
    "pdp3_set_strict_radial_fix": (_I, [_P, _I]),
    "pdp3_gauss_residual": (_I, [_P, _D, _D]),
    "pdp3_gauss_residual_noalloc": (_I, [_P, _D, _D]),
    "pdp3_temperature": (_I, [_P, _I, _D]),
    "pdp3_density": (_I, [_P, _I, _D]),
    "pdp3_bunch_add_loaded": (_I, [_P, _I, _F, _I, _D, _D, _D, _D, _D, _D, C.POINTER(_I)]),
    "pdp3_adjoint_run": (_I, [_P, _I, _I, _I, _F, _F, _D, C.POINTER(_D)]),
    "pdp3_objective_run": (_I, [_P, _I, _I, _F, _F, _D]),
    "pdp3_tangent_run": (_I, [_P, _I, _I, _F, _F, C.POINTER(_D), _D, _D]),
    "pdp3_adjoint_tangent_run": (_I, [_P, _I, _I, _I, _F, _F, C.POINTER(_D), _D, C.POINTER(_D), _D]),
    "pdp3_adjoint_run_batch": (_I, [C.POINTER(_P), _I, _I, _I, _I, _F, _F, _D, C.POINTER(C.POINTER(_D)),
                                    C.POINTER(_I)]),
    "pdp3_ensemble_create": (_I, [C.POINTER(_P), _I, C.POINTER(_P)]),
    "pdp3_ensemble_destroy": (_I, [_P]),
    "pdp3_ensemble_last_error": (C.c_char_p, [_P]),
    "pdp3_ensemble_step": (_I, [_P, _I]),
    "pdp3_ensemble_set_sort_interval": (_I, [_P, _I]),
    "pdp3_ensemble_timing": (_I, [_P, _D, C.POINTER(C.c_longlong)]),
    "pdp3_ensemble_set_tile": (_I, [_P, _I, _I]),
    "pdp3_ensemble_get_tile": (_I, [_P, C.POINTER(_I), C.POINTER(_I)]),
    "pdp3_ensemble_autotune": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I)]),
    "pdp3_set_tile": (_I, [_P, _I, _I]),
    "pdp3_get_tile": (_I, [_P, C.POINTER(_I), C.POINTER(_I)]),
    "pdp3_autotune": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I)]),
    "pdp3_profile_enable": (_I, [_P, _I]),
    "pdp3_profile_get": (_I, [_P, _I, _D, C.POINTER(C.c_longlong)]),
    "pdp3_profile_reset": (_I, [_P]),
    "pdp3_version": (C.c_char_p, []),
}

_lib = None


class Pdp3Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pdp3 error {code}: {msg}")
        self.code = code


def lib():
    """Load libpdp3_hip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m pdp3_amd.build` (HIP library required)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Simulation:
    """One pdp3 context: geometry, fields, species and time on the GPU."""

    def __init__(self, n_grid_r, n_grid_z, r_size, z_size, dt, pusher=PUSHER_BORIS,
                 precision=PREC_FP64, current_time=0.0, end_time=1e300):
        self.L = lib()
        h = _P()
        st = self.L.pdp3_ctx_create(int(n_grid_r), int(n_grid_z), float(r_size), float(z_size), float(current_time),
                                    0.0, 0.0, float(end_time), float(dt), int(pusher), int(precision), C.byref(h))
        if st != OK:
            raise Pdp3Error(st, "pdp3_ctx_create failed")
        self.h = h
        self.nr, self.nz = int(n_grid_r), int(n_grid_z)
        g = np.empty(6)
        self.L.pdp3_geometry_get(self.h, _dp(g))
        self.dr, self.dz = float(g[0]), float(g[1])
        self.dt = float(dt)
        self.precision = precision

    def close(self):
        if getattr(self, "h", None):
            self.L.pdp3_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        if st != OK:
            raise Pdp3Error(st, self.L.pdp3_last_error(self.h).decode())
        return st

    # ---- geometry / fields ----
    def set_pml(self, l1, l2, l3, s1, s2):
        self._chk(self.L.pdp3_geometry_set_pml(self.h, l1, l2, l3, s1, s2))

    def boundary_conditions(self):
        self._chk(self.L.pdp3_efield_boundary_conditions(self.h))

    @property
    def time(self):
        t = C.c_double()
        self.L.pdp3_time_get(self.h, C.byref(t), None)
        return t.value

    def field(self, name):
        comp = COMPONENTS[name] if isinstance(name, str) else int(name)
        out = np.empty((self.nr, self.nz))
        self._chk(self.L.pdp3_grid_download(self.h, comp, _dp(out)))
        return out

    def set_field(self, name, arr):
        comp = COMPONENTS[name] if isinstance(name, str) else int(name)
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.nr, self.nz)
        self._chk(self.L.pdp3_grid_upload(self.h, comp, _dp(a)))

    def h_calc(self):
        self._chk(self.L.pdp3_hfield_calc(self.h))

    def e_calc(self):
        self._chk(self.L.pdp3_efield_calc(self.h))

    def reset_j(self):
        self._chk(self.L.pdp3_current_reset(self.h))

    def get_e(self, r, z):
        r, z = _f64(np.atleast_1d(r)), _f64(np.atleast_1d(z))
        out = np.empty((len(r), 3))
        self._chk(self.L.pdp3_efield_get_field(self.h, _dp(r), _dp(z), len(r), _dp(out)))
        return out

    def get_h(self, r, z):
        r, z = _f64(np.atleast_1d(r)), _f64(np.atleast_1d(z))
        out = np.empty((len(r), 3))
        self._chk(self.L.pdp3_hfield_get_field(self.h, _dp(r), _dp(z), len(r), _dp(out)))
        return out

    # ---- species ----
    def add_species(self, charge_e, mass_e, number):
        idx = _I()
        self._chk(self.L.pdp3_species_add(self.h, int(charge_e), float(mass_e), int(number), C.byref(idx)))
        self._nsp = max(getattr(self, '_nsp', 0), idx.value + 1)
        return idx.value

    def add_plasma(self, charge_e, mass_e, r, z, vr, vphi, vz, w, alive=None):
        idx = self.add_species(charge_e, mass_e, len(r))
        self.upload(idx, r=r, z=z, vr=vr, vphi=vphi, vz=vz, w=w,
                    alive=np.ones(len(r), np.uint8) if alive is None else alive)
        return idx

    def add_bunch_loaded(self, charge_e, mass_e, r, z, vr, vphi, vz, w):
        """A bunch already inside the grid (alive, no injection)."""
        arrs = [_f64(a) for a in (r, z, vr, vphi, vz, w)]
        idx = _I()
        self._chk(self.L.pdp3_bunch_add_loaded(self.h, int(charge_e), float(mass_e), len(arrs[0]),
                                               *(_dp(a) for a in arrs), C.byref(idx)))
        self._nsp = max(getattr(self, '_nsp', 0), idx.value + 1)
        return idx.value

    def add_bunch(self, charge_e, mass_e, r, w, inject_u, duration, hole_duration, velocity, bunch_number=0):
        r, w, u = _f64(r), _f64(w), _f64(inject_u)
        idx = _I()
        self._chk(self.L.pdp3_bunch_add(self.h, int(charge_e), float(mass_e), len(r), _dp(r), _dp(w), _dp(u),
                                        float(duration), float(hole_duration), float(velocity), int(bunch_number),
                                        C.byref(idx)))
        self._nsp = max(getattr(self, '_nsp', 0), idx.value + 1)
        return idx.value

    def init_uniform(self, idx, seed, ppc, density, ramp_cells, temperature_ev):
        self._chk(self.L.pdp3_species_init_uniform(self.h, idx, C.c_uint64(seed), int(ppc), float(density),
                                                   int(ramp_cells), float(temperature_ev)))

    def n_species(self):
        return getattr(self, "_nsp", 0)

    def count(self, idx):
        n = _I()
        self._chk(self.L.pdp3_species_count(self.h, idx, C.byref(n)))
        return n.value

    def upload(self, idx, r=None, z=None, vr=None, vphi=None, vz=None, w=None, alive=None):
        arrs = [_f64(a) for a in (r, z, vr, vphi, vz, w)]
        al = None if alive is None else np.ascontiguousarray(alive, dtype=np.uint8)
        self._chk(self.L.pdp3_species_upload(self.h, idx, *(_dp(a) for a in arrs),
                                             None if al is None else al.ctypes.data_as(_U8)))

    def download(self, idx):
        n = self.count(idx)
        out = {k: np.empty(n) for k in ("r", "z", "vr", "vphi", "vz", "w")}
        alive = np.empty(n, np.uint8)
        self._chk(self.L.pdp3_species_download(self.h, idx, *(_dp(out[k]) for k in ("r", "z", "vr", "vphi", "vz", "w")),
                                               alive.ctypes.data_as(_U8)))
        out["alive"] = alive
        return out

    # ---- stepping ----
    def step(self, n=1):
        self._chk(self.L.pdp3_step(self.h, int(n)))

    def step_v(self, idx):
        self._chk(self.L.pdp3_particles_step_v(self.h, idx))

    def full_current_distribution(self, idx):
        self._chk(self.L.pdp3_particles_full_current_distribution(self.h, idx))

    def charge_weighting(self, idx):
        self._chk(self.L.pdp3_particles_charge_weighting(self.h, idx))

    def gauss_residual(self):
        """(R, rho): R = div E - rho/eps0 on the nodes and the charge density of
        all species, both (n_r, n_z)."""
        R = np.empty((self.nr, self.nz))
        rho = np.empty((self.nr, self.nz))
        self._chk(self.L.pdp3_gauss_residual(self.h, R.ctypes.data_as(_D), rho.ctypes.data_as(_D)))
        return R, rho

    def gauss_residual_noalloc(self):
        """gauss_residual computed in the library's fixed device array (no
        device allocation; grids of at most NOALLOC_MAX_NODES nodes)."""
        R = np.empty((self.nr, self.nz))
        rho = np.empty((self.nr, self.nz))
        self._chk(self.L.pdp3_gauss_residual_noalloc(self.h, R.ctypes.data_as(_D), rho.ctypes.data_as(_D)))
        return R, rho

    def temperature(self, idx):
        """TemperatureClassic of species idx on the device: (n_r, n_z) in eV"""
        out = np.empty((self.nr, self.nz))
        self._chk(self.L.pdp3_temperature(self.h, int(idx), _dp(out)))
        return out

    def density(self, idx):
        """ParticlesDensity::calc_density of species idx on the device: (n_r, n_z)"""
        out = np.empty((self.nr, self.nz))
        self._chk(self.L.pdp3_density(self.h, int(idx), _dp(out)))
        return out

    def adjoint_run(self, nsteps, checkpoint_every, window_steps, z_lo, z_hi):
        """(phi, [dphi/dw per species]) of the axis E_z^2 objective over the
        last window_steps of nsteps iterations (the simulation advances)."""
        counts = [self.count(j) for j in range(self.n_species())]
        outs = [np.zeros(max(n, 1)) for n in counts]
        arr = (_D * len(outs))(*[_dp(o) for o in outs])
        phi = C.c_double()
        self._chk(self.L.pdp3_adjoint_run(self.h, int(nsteps), int(checkpoint_every), int(window_steps),
                                          float(z_lo), float(z_hi), C.byref(phi), arr))
        return phi.value, [o[:n] for o, n in zip(outs, counts)]

    def tangent_run(self, nsteps, window_steps, z_lo, z_hi, dw):
        """(phi, dphi[TANGENT_DIRS]) of the axis E_z^2 objective and its
        directional derivatives along weight tangents dw: one (TANGENT_DIRS, n)
        array (or None) per species (the simulation advances)."""
        arrs = []
        for j in range(self.n_species()):
            a = None if dw is None or j >= len(dw) or dw[j] is None else \
                np.ascontiguousarray(dw[j], dtype=np.float64).reshape(TANGENT_DIRS, self.count(j))
            arrs.append(a)
        ptrs = (_D * max(len(arrs), 1))(*[_dp(a) for a in arrs])
        phi = C.c_double()
        dphi = np.zeros(TANGENT_DIRS)
        self._chk(self.L.pdp3_tangent_run(self.h, int(nsteps), int(window_steps), float(z_lo), float(z_hi), ptrs,
                                          C.byref(phi), _dp(dphi)))
        return phi.value, dphi

    def adjoint_tangent_run(self, nsteps, checkpoint_every, window_steps, z_lo, z_hi, dw):
        """adjoint_run plus the tangent_run derivatives along dw, both around
        the trajectory of the adjoint's own forward sweep: (phi, [dphi/dw per
        species], dphi[TANGENT_DIRS])"""
        counts = [self.count(j) for j in range(self.n_species())]
        arrs = [None if dw[j] is None else np.ascontiguousarray(dw[j], dtype=np.float64).reshape(TANGENT_DIRS, n)
                for j, n in enumerate(counts)]
        pdw = (_D * max(len(arrs), 1))(*[_dp(a) for a in arrs])
        outs = [np.zeros(max(n, 1)) for n in counts]
        pout = (_D * max(len(outs), 1))(*[_dp(o) for o in outs])
        phi = C.c_double()
        dphi = np.zeros(TANGENT_DIRS)
        self._chk(self.L.pdp3_adjoint_tangent_run(self.h, int(nsteps), int(checkpoint_every), int(window_steps),
                                                  float(z_lo), float(z_hi), pdw, C.byref(phi), pout, _dp(dphi)))
        return phi.value, [o[:n] for o, n in zip(outs, counts)], dphi

    def objective_run(self, nsteps, window_steps, z_lo, z_hi):
        phi = C.c_double()
        self._chk(self.L.pdp3_objective_run(self.h, int(nsteps), int(window_steps), float(z_lo), float(z_hi),
                                            C.byref(phi)))
        return phi.value

    def set_tile(self, ppt, waves):
        self._chk(self.L.pdp3_set_tile(self.h, int(ppt), int(waves)))

    def tile(self):
        a, b = _I(), _I()
        self._chk(self.L.pdp3_get_tile(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune(self, steps_per_candidate=2):
        """Times real steps with every tile shape, keeps the fastest; returns (ppt, waves)."""
        a, b = _I(), _I()
        self._chk(self.L.pdp3_autotune(self.h, int(steps_per_candidate), C.byref(a), C.byref(b)))
        return a.value, b.value

    def sync(self):
        self._chk(self.L.pdp3_sync(self.h))

    def sort(self):
        self._chk(self.L.pdp3_sort_particles(self.h))

    def set_sort_interval(self, n):
        self._chk(self.L.pdp3_set_sort_interval(self.h, int(n)))

    def set_strict_radial_fix(self, on=True):
        """FP64 plain form: evaluate strictly radial face crossings without
        the reference's operator-precedence slip (analysis option)"""
        self._chk(self.L.pdp3_set_strict_radial_fix(self.h, 1 if on else 0))

    # ---- timing ----
    def profile(self, enable=True):
        self._chk(self.L.pdp3_profile_enable(self.h, 1 if enable else 0))

    def profile_reset(self):
        self._chk(self.L.pdp3_profile_reset(self.h))

    def profile_get(self, kernel):
        ms, n = C.c_double(), C.c_longlong()
        self._chk(self.L.pdp3_profile_get(self.h, kernel, C.byref(ms), C.byref(n)))
        return ms.value, n.value


def adjoint_run_batch(sims, nsteps, checkpoint_every, window_steps, z_lo, z_hi):
    """Simulation.adjoint_run over independent members at once (one host
    thread and stream per member): ([phi per member], [[dphi/dw per species]
    per member]); every simulation advances."""
    L = lib()
    n = len(sims)
    handles = (_P * n)(*[s.h for s in sims])
    counts = [[s.count(j) for j in range(s.n_species())] for s in sims]
    outs = [[np.zeros(max(c, 1)) for c in cs] for cs in counts]
    rows = [(_D * max(len(o), 1))(*[_dp(a) for a in o]) for o in outs]
    table = (C.POINTER(_D) * n)(*[C.cast(r, C.POINTER(_D)) for r in rows])
    phi = np.zeros(n)
    status = (_I * n)()
    st = L.pdp3_adjoint_run_batch(handles, n, int(nsteps), int(checkpoint_every), int(window_steps), float(z_lo),
                                  float(z_hi), _dp(phi), table, status)
    if st != OK:
        bad = [m for m in range(n) if status[m] != OK]
        msg = "; ".join(f"member {m}: {L.pdp3_last_error(sims[m].h).decode()}" for m in bad)
        raise Pdp3Error(st, msg)
    return list(phi), [[a[:c] for a, c in zip(o, cs)] for o, cs in zip(outs, counts)]


class Ensemble:
    """Independent FP32-form simulations advanced together (one graph replay
    per step with one particle launch over every instance)."""

    def __init__(self, sims):
        self.L = lib()
        self.sims = list(sims)
        arr = (_P * len(self.sims))(*[s.h for s in self.sims])
        h = _P()
        st = self.L.pdp3_ensemble_create(arr, len(self.sims), C.byref(h))
        if st != 0:
            raise Pdp3Error(st, "pdp3_ensemble_create failed (FP32 contexts with one common dt)")
        self.h = h

    def _chk(self, st):
        if st != 0:
            raise Pdp3Error(st, self.L.pdp3_ensemble_last_error(self.h).decode())

    def step(self, n=1):
        self._chk(self.L.pdp3_ensemble_step(self.h, int(n)))

    def set_sort_interval(self, n):
        self._chk(self.L.pdp3_ensemble_set_sort_interval(self.h, int(n)))

    def timing(self):
        """(device ms of the stepped graphs, steps) since the last call"""
        ms, n = C.c_double(), C.c_longlong()
        self._chk(self.L.pdp3_ensemble_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_tile(self, ppt, waves):
        self._chk(self.L.pdp3_ensemble_set_tile(self.h, int(ppt), int(waves)))

    def tile(self):
        a, b = _I(), _I()
        self._chk(self.L.pdp3_ensemble_get_tile(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune(self, steps_per_candidate=2):
        a, b = _I(), _I()
        self._chk(self.L.pdp3_ensemble_autotune(self.h, int(steps_per_candidate), C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if getattr(self, "h", None):
            self.L.pdp3_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


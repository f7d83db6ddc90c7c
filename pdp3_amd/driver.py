"""ctypes binding of libpdp3_host.so (include/pdp3_host.h): the TESTMODE
initial loaders, the parameters.xml reader and the pdp3 driver
(LoadInitParam, src/loadInitParam.cpp) run in-process.

The command-line driver is pdp3_amd/bin/pdp3 (``pdp3 -f parameters.xml``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(HERE, "lib", "libpdp3_host.so")
DRIVER = os.path.join(HERE, "bin", "pdp3")

_D = C.POINTER(C.c_double)
SIGNATURES = {
    "pdp3_host_random_reverse": (C.c_double, [C.c_double, C.c_int]),
    "pdp3_host_load_cylindrical": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                             C.c_double, C.c_double, _D, _D, _D]),
    "pdp3_host_maxwell_juttner": (C.c_int, [C.c_int, C.c_double, _D]),
    "pdp3_host_velocity_distribution": (C.c_int, [C.c_int, C.c_double, C.c_double, _D, _D, _D]),
    "pdp3_host_bunch_positions": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_double, _D, _D]),
    "pdp3_host_bunch_inject_draws": (C.c_int, [C.c_int, _D]),
    "pdp3_host_mb_table": (C.c_int, [_D]),
    "pdp3_host_mj_table": (C.c_int, [_D]),
    "pdp3_host_parameters_dump": (C.c_int, [C.c_char_p, C.c_char_p, C.c_size_t]),
    "pdp3_host_run": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int)]),
    "pdp3_host_last_error": (C.c_char_p, []),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB):
            raise ImportError(f"{HOST_LIB} is not built: run `python -m pdp3_amd.build`")
        L = C.CDLL(HOST_LIB)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


class DriverError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (status {code})")
        self.code = code


def _ptr(a):
    return a.ctypes.data_as(_D)


def _check(st):
    if st != 0:
        raise DriverError(st, lib().pdp3_host_last_error().decode())


def random_reverse(vel, power):
    return lib().pdp3_host_random_reverse(float(vel), int(power))


def load_cylindrical(n, r_size, z_size, dr, dz, n1, n2, left=0.0):
    r, z, w = (np.empty(n) for _ in range(3))
    _check(lib().pdp3_host_load_cylindrical(n, r_size, z_size, dr, dz, n1, n2, left, _ptr(r), _ptr(z), _ptr(w)))
    return r, z, w


def maxwell_juttner(n, temperature):
    e = np.empty(max(n, 1))
    _check(lib().pdp3_host_maxwell_juttner(n, temperature, _ptr(e)))
    return e[:n]


def velocity_distribution(n, mass_kg, temperature):
    vr, vphi, vz = (np.empty(max(n, 1)) for _ in range(3))
    _check(lib().pdp3_host_velocity_distribution(n, mass_kg, temperature, _ptr(vr), _ptr(vphi), _ptr(vz)))
    return vr[:n], vphi[:n], vz[:n]


def bunch_positions(n, radius, dr, dz, duration, velocity, density):
    r, w = np.empty(n), np.empty(n)
    _check(lib().pdp3_host_bunch_positions(n, radius, dr, dz, duration, velocity, density, _ptr(r), _ptr(w)))
    return r, w


def bunch_inject_draws(n):
    u = np.empty(max(n, 1))
    _check(lib().pdp3_host_bunch_inject_draws(n, _ptr(u)))
    return u[:n]


def mb_table():
    t = np.empty(1000)
    _check(lib().pdp3_host_mb_table(_ptr(t)))
    return t


def mj_table():
    t = np.empty(1000)
    _check(lib().pdp3_host_mj_table(_ptr(t)))
    return t


def parameters(path):
    """Parameters(path) as a dict of strings (the reader's own view)."""
    buf = C.create_string_buffer(1 << 16)
    _check(lib().pdp3_host_parameters_dump(path.encode(), buf, len(buf)))
    out = {}
    for line in buf.value.decode().splitlines():
        k, _, v = line.partition("=")
        out[k] = v
    return out


def run(path, quiet=True):
    """LoadInitParam(path).run() in this process; returns the step count."""
    steps = C.c_int(0)
    _check(lib().pdp3_host_run(path.encode(), 1 if quiet else 0, C.byref(steps)))
    return steps.value

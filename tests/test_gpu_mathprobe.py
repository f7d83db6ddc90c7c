"""How often the GPU's FP64 pow(x,-0.5) and log differ from the host's glibc,
and that the correctly rounded replacements used by the FP64 plain form
(pdp3_amd/csrc/crmath.hpp) agree with the exact value (mpmath, 200 bits).

glibc itself is correctly rounded in all but ~0.04% of these inputs, so the
replacements must agree with glibc on >= 99.9% of them."""
import ctypes
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _inputs(rng, n):
    # gamma_inv arguments 1 + (u/c)^2 and deposit log arguments 1 + dr/r
    a = 1.0 + (10 ** rng.uniform(-8, 1, n // 2))
    x = rng.choice([-1, 1], n - n // 2) * 10 ** rng.uniform(-12, np.log10(0.5), n - n // 2)
    return np.concatenate([a, 1.0 + x])


def test_probe_mismatch_rates():
    lib = ctypes.CDLL(os.path.join(HERE, "native", "libmathprobe.so"))
    rng = np.random.default_rng(11)
    x = _inputs(rng, 200000)
    out = np.empty((len(x), 5))
    assert lib.mathprobe_run(x.ctypes.data_as(ctypes.c_void_p), len(x), out.ctypes.data_as(ctypes.c_void_p)) == 0
    g_rs = np.array([math.pow(v, -0.5) for v in x])
    g_lg = np.array([math.log(v) for v in x])
    g_sq = np.sqrt(x)
    rates = {
        "ocml pow(x,-0.5)": np.mean(out[:, 0] != g_rs),
        "ocml log": np.mean(out[:, 1] != g_lg),
        "sqrt": np.mean(out[:, 2] != g_sq),
        "rsqrt_cr": np.mean(out[:, 3] != g_rs),
        "log_cr": np.mean(out[:, 4] != g_lg),
    }
    print({k: f"{v:.5%}" for k, v in rates.items()})
    assert rates["sqrt"] == 0.0
    assert rates["rsqrt_cr"] <= 1e-3
    assert rates["log_cr"] <= 1e-3
    # spot-check against the exact value where the two disagree
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.prec = 200
    for j in np.nonzero(out[:, 3] != g_rs)[0][:50]:
        assert out[j, 3] == float(1 / mpmath.sqrt(mpmath.mpf(float(x[j]))))
    for j in np.nonzero(out[:, 4] != g_lg)[0][:50]:
        assert out[j, 4] == float(mpmath.log(mpmath.mpf(float(x[j]))))

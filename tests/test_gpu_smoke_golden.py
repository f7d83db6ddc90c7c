"""The reference's own end-to-end smoke case replayed on the GPU.

Initial particles come from the oracle's TESTMODE loaders (exactly the
reference's), the loop runs through pdp3_step, and the fifth probe frame is
compared with the reference's goldens (6 printed digits) and with the oracle
(normwise, FP64 bar 1e-10).
"""
import numpy as np
import pytest

import smoke_case as sc
from helpers import normwise

pytestmark = pytest.mark.gpu


def test_smoke_goldens_on_gpu(orc):
    from pdp3_amd import capi

    plasma, bunch = sc.initial_state(orc)
    g, T = sc.GEOM, sc.TIME
    gpu = capi.Simulation(g["n_grid_r"], g["n_grid_z"], g["r_size"], g["z_size"], T["dt"],
                          current_time=T["current_time"], end_time=T["end_time"])
    ora = orc.Sim(g["n_grid_r"], g["n_grid_z"], g["r_size"], g["z_size"], T["dt"],
                  current_time=T["current_time"], end_time=T["end_time"])
    gpu.boundary_conditions()
    for p in plasma:
        gpu.add_plasma(p["charge"], p["mass"], p["r"], p["z"], p["vr"], p["vphi"], p["vz"], p["w"])
        ora.add_plasma(p["charge"], p["mass"], p["r"], p["z"], p["vr"], p["vphi"], p["vz"], p["w"])
    b = bunch
    gpu.add_bunch(b["charge"], b["mass"], b["r"], b["w"], b["inject_u"], b["duration"], b["hole_duration"],
                  b["velocity"], b["bunch_number"])
    ora.add_bunch(b["charge"], b["mass"], b["r"], b["w"], b["inject_u"], b["duration"], b["hole_duration"],
                  b["velocity"], b["bunch_number"])
    dumps, it, checked = 0, 0, False
    while gpu.time < T["end_time"]:
        fires = sc.probe_fires(gpu.time, T["dt"])
        gpu.step(1)
        ora.step(1)
        if fires:
            if dumps == sc.FRAME_NO:
                assert it == 20
                for comp in sc.FRAMES:
                    name = "rho_bunch" if comp == "rho_beam" else comp
                    ours = sc.frame(gpu.field(name), comp)
                    rel_max, small_max, _ = sc.compare(ours, sc.golden(comp))
                    assert rel_max < 2e-5 and small_max < 1e-9, (comp, rel_max, small_max)
                    e = normwise(gpu.field(name), ora.field(name))
                    print(comp, "golden rel", rel_max, "oracle normwise", e)
                    assert e <= 1e-10, (comp, e)
                checked = True
            dumps += 1
        it += 1
    assert checked and it == 23

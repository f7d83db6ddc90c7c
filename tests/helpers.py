"""Shared helpers for the parity tests: mirror a GPU simulation's initial
state into the oracle and compare states normwise."""
import numpy as np

FIELDS = ("E_r", "E_phi", "E_z", "H_r", "H_phi", "H_z", "HH_r", "HH_phi", "HH_z", "J_r", "J_phi", "J_z", "rho_bunch")


def normwise(ours, ref):
    """max |ours - ref| / max |ref| (0 when both are zero)."""
    ours = np.asarray(ours, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.max(np.abs(ref)) if ref.size else 0.0
    num = np.max(np.abs(ours - ref)) if ref.size else 0.0
    if den == 0.0:
        return 0.0 if num == 0.0 else np.inf
    return float(num / den)


def oracle_from_workload(orc, w, gpu_sim, plasma_states):
    """An oracle Sim with the workload's geometry, PML, pusher and the given
    plasma states (downloaded from the GPU) plus the workload's bunches."""
    sim = orc.Sim(w.n_grid_r, w.n_grid_z, w.r_size, w.z_size, w.dt, pusher=w.pusher)
    pml = w.pml()
    if pml:
        sim.set_pml(*pml)
    for p, st in zip(w.plasma, plasma_states):
        sim.add_plasma(p.charge, p.mass, st["r"], st["z"], st["vr"], st["vphi"], st["vz"], st["w"], st["alive"])
    for b in w.bunches:
        sim.add_bunch(b.charge, b.mass, b.r, b.w, b.inject_u, b.duration, b.hole_duration, b.velocity, b.bunch_number)
    return sim


def compare_fields(gpu, ora, names=FIELDS):
    return {n: normwise(gpu.field(n), ora.field(n)) for n in names}


def compare_species(gpu_state, ora_state):
    """normwise errors of positions and momenta of alive particles, plus the
    count of alive-flag mismatches"""
    ga, oa = gpu_state["alive"].astype(bool), ora_state["alive"].astype(bool)
    mism = int(np.sum(ga != oa))
    m = ga & oa
    out = {"alive_mismatch": mism}
    for k in ("r", "z", "vr", "vphi", "vz"):
        out[k] = normwise(gpu_state[k][m], ora_state[k][m]) if m.any() else 0.0
    return out

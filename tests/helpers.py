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


EPSILON0 = 8.85e-12  # include/constant.h:12


def gauss_residual_np(er, ez, rho, dr, dz):
    """div E - rho/eps0 on the nodes (r_i, z_k) where the charge weighting puts
    rho, with E_r at (r_{i+1/2}, z_k) and E_z at (r_i, z_{k+1/2}): the
    divergence written out from the Gauss theorem on the node's ring cell
    (radial faces at r_{i-1/2}, r_{i+1/2}; the axis cell is the disc of radius
    dr/2), independently of the device kernel.  0 on k = 0, k = nz-1 and
    i = nr-1."""
    nr, nz = er.shape
    R = np.zeros((nr, nz))
    for i in range(nr - 1):
        if i == 0:
            area_out = 2 * np.pi * (dr / 2) * dz        # side of the disc
            vol = np.pi * (dr / 2) ** 2 * dz
            flux_r = area_out * er[0, 1:-1]
            cap = np.pi * (dr / 2) ** 2                  # top / bottom of the disc
        else:
            a_out = 2 * np.pi * (i + 0.5) * dr * dz
            a_in = 2 * np.pi * (i - 0.5) * dr * dz
            vol = 2 * np.pi * i * dr * dr * dz
            flux_r = a_out * er[i, 1:-1] - a_in * er[i - 1, 1:-1]
            cap = 2 * np.pi * i * dr * dr
        flux_z = cap * (ez[i, 1:-1] - ez[i, :-2])
        R[i, 1:-1] = (flux_r + flux_z) / vol - rho[i, 1:-1] / EPSILON0
    return R


def oracle_rho_all(orc, sim):
    """charge density of every species of an oracle Sim (the reference's
    charge weighting, src/particles.cpp:101-226)"""
    import ctypes as C
    out = np.zeros((sim.nr, sim.nz))
    for idx in sim.species:
        st = orc.lib().or_sim_charge_weighting(sim.h, idx, out.ctypes.data_as(C.POINTER(C.c_double)))
        assert st == 0
    return out

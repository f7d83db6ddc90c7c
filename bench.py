"""Benchmark of the r-z PIC step (push + deposit + FDTD) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config cfg3]

Workload (default cfg3): 512 x 8192 cells, dr = dz = 2.5e-4 m,
dt = 0.5 dr / c, PML 16 cells, electrons at 100 per cell (4.19e8) with a
density ramping linearly from 0 to 1e18 m^-3 over the first 1000 z-cells,
immobile ion background, Gaussian bunch (sigma_r 0.5 mm, sigma_z 1 mm, peak
5e17 m^-3, 0.999 c, 1e6 macro-particles), Boris pusher.  Synthetic data
(device Philox loader, seeded).

Both precision forms run the same workload from the same initial state:
the FP32 cell-relative form (the reported value) and the FP64 plain form
(reference arithmetic).  One JSON line goes to stdout with particle-steps/s
of the FP32 form, the FP64 form's rate, E_z error of FP32 against FP64, the
roofline of the dominant kernel (timed live with HIP events on the
library's stream) and the CPU oracle timed on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s (push+deposit+FDTD), % HBM peak; FP32 vs FP64 at E_z error"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per particle-step of the fused particle kernel: FP32 form
# reads cell 4 + offsets 8 + u 12 + w 4 and writes 24; FP64 plain form reads
# r,z 16 + v 24 + w 8 + alive 1 and writes 40 + 1
BYTES_PER_PARTICLE = {"fp32": 52.0, "fp64": 90.0}
# per cell and launch: gather of E and H-half (6 x 8 B) and the J flush (3 x 8 B)
BYTES_PER_CELL = 72.0


AUTOTUNE_STEPS = 12  # 6 tile shapes x 2 steps


def run_form(wl, capi, w, precision, steps, warmup, sort_interval):
    sim = wl.build_simulation(w, precision=precision, sort_interval=sort_interval)
    sim.sync()
    sim.step(warmup)
    sim.sync()
    # the library's autotuner picks the tile shape of the FP32 particle kernel
    # (AUTOTUNE_STEPS real steps: 2 per shape); the FP64 form has one shape and
    # takes the same steps, so that both forms are compared at one time
    if precision == capi.PREC_FP32:
        tile = list(sim.autotune(AUTOTUNE_STEPS // 6))
    else:
        tile = None
        sim.step(AUTOTUNE_STEPS)
    sim.sync()
    sim.profile(True)
    sim.profile_reset()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    t1 = time.perf_counter()
    ms, launches = sim.profile_get(capi.KERNEL_PARTICLES)
    ms_b, _ = sim.profile_get(capi.KERNEL_BUNCH)
    ms_h, _ = sim.profile_get(capi.KERNEL_H)
    ms_e, _ = sim.profile_get(capi.KERNEL_E)
    ms_s, ls = sim.profile_get(capi.KERNEL_SORT)
    sim.profile(False)
    ez = sim.field("E_z")
    out = dict(wall_s=t1 - t0, particle_ms=ms, particle_launches=launches, bunch_ms=ms_b, h_ms=ms_h, e_ms=ms_e,
               sort_ms=ms_s, sort_launches=ls, tile=tile)
    sim.close()
    return out, ez


def ez_error_full(wl, capi, w, steps, sort_interval):
    """E_z of the FP32 form against the FP64 form on the full grid after
    `steps` steps from the same initial state, all runs with the same periodic
    cell sort (the production loop).  The reference for the precision figure
    is the FP64 form with the strict-radial precedence fix
    (pdp3_set_strict_radial_fix): the reference's strict_motion_weighting
    parses r0*r0/i_n*dr as ((r0^2)/i_n) dr, which injects a spurious J_r
    whenever a particle with |dz| < 1e-15 m per step crosses a radial face
    (about one such event per 1000 cfg3 steps); the FP32 form's closed forms
    have no such term.  The comparison with the plain FP64 form (the
    reference's arithmetic, events included) is reported beside it."""
    runs = [(capi.PREC_FP32, False), (capi.PREC_FP64, True), (capi.PREC_FP64, False)]
    ez = []
    for p, fix in runs:
        sim = wl.build_simulation(w, precision=p, sort_interval=sort_interval)
        if fix:
            sim.set_strict_radial_fix(True)
        sim.step(steps)
        ez.append(sim.field("E_z"))
        sim.close()

    def cmp(a, b):
        d = a - b
        return (float(np.sqrt(np.mean(d ** 2) / np.mean(b ** 2))), float(np.max(np.abs(d)) / np.max(np.abs(b))))

    rms, mx = cmp(ez[0], ez[1])
    rms_p, mx_p = cmp(ez[0], ez[2])
    rms_pp, mx_pp = cmp(ez[1], ez[2])
    return dict(E_z_rel_rms=rms, E_z_rel_max=mx, after_steps=steps,
                reference="FP64 form with the strict-radial precedence fix",
                vs_fp64_plain=dict(E_z_rel_rms=rms_p, E_z_rel_max=mx_p,
                                   fp64_fixed_vs_plain_rel_rms=rms_pp, fp64_fixed_vs_plain_rel_max=mx_pp),
                sample=f"full {w.cells_r}x{w.cells_z} grid, cell sort every {sort_interval} steps")


def momentum_error(wl, capi, name, cells_z, steps):
    """Particle momentum (gamma v) and E_z errors of the FP32 form against the
    FP64 plain form on a z-slab of the workload: the same spacing, ppc, density
    profile and bunch over the first cells_z cells, no cell sort (both runs keep
    one particle order, so particles pair up by index)."""
    w = wl.config(name)
    w.cells_z = cells_z
    sims = [wl.build_simulation(w, precision=p) for p in (capi.PREC_FP32, capi.PREC_FP64)]
    sims[1].set_strict_radial_fix(True)  # the precision comparison: see ez_error_full
    for s_ in sims:
        s_.step(steps)
    ez = [s_.field("E_z") for s_ in sims]
    mom = []
    for s_ in sims:
        d = s_.download(0)
        g = 1.0 / np.sqrt(1.0 - (d["vr"] ** 2 + d["vphi"] ** 2 + d["vz"] ** 2) / wl.LIGHT_SPEED ** 2)
        mom.append(np.stack([g * d["vr"], g * d["vphi"], g * d["vz"]]))
        del d
        s_.close()
    dp = mom[0] - mom[1]
    # per particle |dp| / rms |p|; the largest come from particles whose wall or
    # axis reflection fell on another step in one of the forms (a sign flip)
    rms_p = np.sqrt(np.mean(np.sum(mom[1] ** 2, axis=0)))
    per = np.sqrt(np.sum(dp ** 2, axis=0)) / rms_p
    return dict(momentum_rel_rms=float(np.sqrt(np.sum(dp ** 2) / np.sum(mom[1] ** 2))),
                momentum_rel_max=float(np.max(np.abs(dp)) / np.max(np.abs(mom[1]))),
                momentum_per_particle_p9999=float(np.quantile(per, 0.9999)),
                particles_above_1e_3=int(np.count_nonzero(per > 1e-3)),
                E_z_rel_rms=float(np.sqrt(np.mean((ez[0] - ez[1]) ** 2) / np.mean(ez[1] ** 2))),
                E_z_rel_max=float(np.max(np.abs(ez[0] - ez[1])) / np.max(np.abs(ez[1]))),
                sample=f"{w.cells_r}x{cells_z}-cell z-slab of {name}, {mom[1].shape[1]} plasma particles, "
                       f"{steps} steps, no cell sort, FP64 form with the strict-radial precedence fix")


def host_cpu():
    """(threads for the CPU baseline, logical CPUs visible, CPU model): the
    process's CPU share, i.e. its affinity set, capped by OMP_NUM_THREADS when
    that is set (one GPU's share of a multi-GPU host)"""
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:
        visible = os.cpu_count() or 1
    threads = visible
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, visible, model


def cpu_baseline(wl, capi, w, threads, visible, model, cells_z=1024, nsteps=3):
    """The oracle (FP64, reference-faithful: push/move loops threaded, deposits
    serial as in the reference) on a bounded sample of the workload: the same
    cells in r, the first cells_z cells in z at the same ppc and spacing
    (uniform density), nsteps full steps (H, push, deposit, E)."""
    from oracle import oracle as orc

    s = wl.config(w.name, seed=w.seed)
    s.cells_z = min(cells_z, w.cells_z)
    s.bunches = []
    for p in s.plasma:
        p.ramp_cells = 0
    sim = wl.build_simulation(s, precision=capi.PREC_FP64)
    states = [sim.download(j) for j in range(len(s.plasma))]
    sim.close()
    ora = orc.Sim(s.n_grid_r, s.n_grid_z, s.r_size, s.z_size, s.dt, pusher=s.pusher, threads=threads)
    pml = s.pml()
    if pml:
        ora.set_pml(*pml)
    for p, st in zip(s.plasma, states):
        ora.add_plasma(p.charge, p.mass, st["r"], st["z"], st["vr"], st["vphi"], st["vz"], st["w"], st["alive"])
    del states
    n = sum(s.plasma_count(p) for p in s.plasma)
    t0 = time.perf_counter()
    ora.step(nsteps)
    dt = time.perf_counter() - t0
    return dict(value=n * nsteps / dt, unit="particle-steps/s", cores=threads, kind="port",
                host={"cpu_model": model, "logical_cpus_visible": visible, "omp_threads": threads},
                sample=f"{s.cells_r}x{s.cells_z}-cell slab of the same grid spacing, {n} particles "
                       f"({s.plasma[0].ppc} ppc), {nsteps} full steps, {dt:.1f} s")


def copy_probe(nbytes):
    """A plain device-to-device copy of nbytes through the HIP runtime, for
    scaling the PMC byte counters (FETCH_SIZE / WRITE_SIZE) of the particle
    kernel: the counters of this copy are compared with 2 x nbytes moved."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes, "launches": 3}))


def describe_workload(w):
    """The workload named by its sizes: grid, species, bunch, PML, pusher."""
    parts = []
    for p in w.plasma:
        kind = "electrons" if p.charge < 0 else f"ions (charge {p.charge}, mass {p.mass:g} m_e)"
        dens = (f"0->{p.density:g} m^-3 ramp over {p.ramp_cells} z-cells" if p.ramp_cells
                else f"{p.density:g} m^-3")
        parts.append(f"{p.ppc} ppc {kind} ({w.plasma_count(p)}), {dens}")
    bunch = ", ".join(f"bunch of {len(b.r)} macro-particles" for b in w.bunches)
    return (f"{w.name}: r-z {w.cells_r}x{w.cells_z} cells, dr=dz={w.dr} m, " + "; ".join(parts)
            + (f"; {bunch}" if bunch else "") + f"; PML {w.pml_cells}, {'Vay' if w.pusher == 1 else 'Boris'}")


def traffic_from_profile(name):
    """HBM bytes per launch of the dominant kernel as measured by rocprofv3
    (FETCH_SIZE and WRITE_SIZE, each scaled by the counter's reading for a
    device-to-device copy of known size), read from the committed summary
    profiles/<name>.md (line `traffic_bytes_per_launch: N`); None when absent."""
    path = os.path.join(ROOT, "profiles", name + ".md")
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1].split()[0])
    except OSError:
        return None
    return None


def bench_ensemble(a):
    """cfg4: 128 independent instances (per-instance grid, density, ppc, bunch
    ratio drawn from the seed), FP32 form, one graph replay per step with one
    particle launch; the tile shape picked by the autotuner."""
    from pdp3_amd import capi
    from pdp3_amd import workloads as wl

    insts = wl.ensemble_config(a.instances)
    n_part = sum(i.particles for i in insts)
    cells = sum(i.cells_r * i.cells_z for i in insts)
    print(f"# ensemble of {len(insts)} instances, {n_part} particles, {cells} cells", file=sys.stderr, flush=True)
    sims = [wl.build_instance(i, injected=not a.loaded_bunches) for i in insts]
    ens = capi.Ensemble(sims)
    ens.set_sort_interval(a.sort_interval)
    tile = ens.autotune(1)
    ens.step(a.warmup)
    ens.timing()
    t0 = time.perf_counter()
    ens.step(a.steps)
    t1 = time.perf_counter()
    dev_ms, n = ens.timing()
    value = n_part * a.steps / (t1 - t0)
    out = {
        "metric": "particle-steps/s (ensemble, FP32 form)", "value": value, "unit": "particle-steps/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * (t1 - t0) / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"cfg4: {len(insts)} r-z instances, r cells in 64/128/256 x z cells in "
                               f"512..4096, 5..40 ppc, densities 1e16..1e18, Gaussian bunches "
                               f"{'loaded at t = 0' if a.loaded_bunches else 'injected over 6 sigma_z / v'}",
                   "particles": n_part, "cells": cells, "tile_ppt_waves": list(tile),
                   "sort_interval": a.sort_interval},
        "device_ms_per_step": dev_ms / max(n, 1),
        "roofline": {"bound": "hbm", "achieved": (n_part * BYTES_PER_PARTICLE["fp32"] + cells * 304.0)
                     / (dev_ms / max(n, 1) * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "note": "whole step (H, J reset, particles, E) against 52 B per particle + 304 B per cell"},
    }
    out["roofline"]["frac"] = out["roofline"]["achieved"] / HBM_PEAK_GBS
    ens.close()
    for s_ in sims:
        s_.close()
    err_steps = 1000 if a.error_steps is None else a.error_steps
    if err_steps > 0:
        out["error_fp32_vs_fp64"] = ensemble_error(wl, capi, insts, err_steps, not a.loaded_bunches)
        print(f"# cfg4 error: {out['error_fp32_vs_fp64']}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


def ensemble_error(wl, capi, insts, steps, injected, sample=8):
    """E_z and momentum errors of the FP32 form against the FP64 plain form
    (strict-radial precedence fix, as in ez_error_full) after `steps` steps,
    on a sample of the ensemble's members, each run alone (an ensemble member
    equals its single run: tests/test_gpu_ensemble.py), no cell sort so that
    particles pair up by index.  Reports the largest figure over the sample."""
    rng = np.random.default_rng(11)
    picks = sorted(rng.choice(len(insts), size=min(sample, len(insts)), replace=False).tolist())
    worst = dict(E_z_rel_rms=0.0, E_z_rel_max=0.0, momentum_rel_rms=0.0, momentum_rel_max=0.0,
                 fp64_perturbed_1e_7_E_z_rel_rms=0.0)
    per = []
    for j in picks:
        ez, mom = [], []
        for run, p in enumerate((capi.PREC_FP32, capi.PREC_FP64, capi.PREC_FP64)):
            sim = wl.build_instance(insts[j], precision=p, injected=injected)
            if p == capi.PREC_FP64:
                sim.set_strict_radial_fix(True)
            if run == 2:
                # the FP64 run again with every initial velocity component
                # perturbed by 1e-7 relative (the size of one FP32 rounding):
                # how far the loop itself carries such a change
                st = sim.download(0)
                prng = np.random.default_rng(5)
                f = lambda: 1.0 + 1e-7 * prng.standard_normal(len(st["vr"]))
                sim.upload(0, st["r"], st["z"], st["vr"] * f(), st["vphi"] * f(), st["vz"] * f(), st["w"], st["alive"])
                del st
            sim.step(steps)
            ez.append(sim.field("E_z"))
            d = sim.download(0)
            g = 1.0 / np.sqrt(1.0 - (d["vr"] ** 2 + d["vphi"] ** 2 + d["vz"] ** 2) / wl.LIGHT_SPEED ** 2)
            mom.append(np.stack([g * d["vr"], g * d["vphi"], g * d["vz"]]))
            sim.close()
        d_ez = ez[0] - ez[1]
        d_p = mom[0] - mom[1]
        fig = dict(E_z_rel_rms=float(np.sqrt(np.mean(d_ez ** 2) / np.mean(ez[1] ** 2))),
                   E_z_rel_max=float(np.max(np.abs(d_ez)) / np.max(np.abs(ez[1]))),
                   momentum_rel_rms=float(np.sqrt(np.sum(d_p ** 2) / np.sum(mom[1] ** 2))),
                   momentum_rel_max=float(np.max(np.abs(d_p)) / np.max(np.abs(mom[1]))))
        d64 = ez[2] - ez[1]
        fig["fp64_perturbed_1e-7_E_z_rel_rms"] = float(np.sqrt(np.mean(d64 ** 2) / np.mean(ez[1] ** 2)))
        fig["fp64_perturbed_1e_7_E_z_rel_rms"] = fig.pop("fp64_perturbed_1e-7_E_z_rel_rms")
        for k in worst:
            worst[k] = max(worst[k], fig[k])
        i = insts[j]
        per.append(dict(member=j, cells=[i.cells_r, i.cells_z], ppc=i.ppc, density=i.density,
                        bunch_ratio=round(i.bunch_ratio, 3), **{k: float("%.3g" % v) for k, v in fig.items()}))
    worst["per_member"] = per
    worst["after_steps"] = steps
    worst["reference"] = "FP64 form with the strict-radial precedence fix"
    worst["sample"] = (f"members {picks} of {len(insts)}, each run alone, no cell sort; the largest figure "
                       f"over the sample; fp64_perturbed_1e_7: the FP64 run against itself with the initial "
                       f"velocities perturbed by 1e-7 relative once")
    return worst


def bench_adjoint(a):
    """cfg5: gradient of the axis E_z^2 objective with respect to 16 bump
    amplitudes and the ramp slope (FP32 forward, checkpoints every 50 steps,
    hand-written adjoint), timed, and checked against central differences of
    the FP64 plain form (relative step 1e-4) on the components in --fd."""
    import torch

    from pdp3_amd import adjoint as adj
    from pdp3_amd import capi

    spec = adj.cfg5_spec(steps=a.adj_steps, window_steps=min(300, a.adj_steps)) if a.adj_steps else adj.cfg5_spec()
    prob = adj.BumpRampProblem(spec)
    th = prob.theta0()
    log = lambda *m: print("#", *m, file=sys.stderr, flush=True)
    log(f"cfg5: {spec.cells_r}x{spec.cells_z} cells, {spec.steps} steps")
    t0 = time.perf_counter()
    phi_fwd = prob.objective(th, capi.PREC_FP32)
    t1 = time.perf_counter()
    log(f"forward only: {t1 - t0:.2f} s, phi {phi_fwd}")
    theta = th.clone().requires_grad_(True)
    phi = adj.PicObjective.apply(theta, prob, capi.PREC_FP32)
    phi.backward()
    t2 = time.perf_counter()
    log(f"gradient: {t2 - t1:.2f} s")
    g = theta.grad.numpy()
    # the same gradient with the FP64 forward (the reference's arithmetic)
    theta64 = th.clone().requires_grad_(True)
    phi64 = adj.PicObjective.apply(theta64, prob, capi.PREC_FP64)
    phi64.backward()
    g64 = theta64.grad.numpy()
    err = {"phi_rel": abs(float(phi) - float(phi64)) / abs(float(phi64)),
           "gradient_normwise": float(np.max(np.abs(g - g64)) / np.max(np.abs(g64))),
           "gradient_max_component_rel": float(np.max(np.abs(g - g64) / np.abs(g64))),
           "reference": "the same adjoint with the FP64 plain-form forward"}
    log(f"FP32 forward against FP64 forward: {err}")
    fd_idx = [int(x) for x in a.fd.split(",")] if a.fd else []
    fd = None
    if fd_idx:
        fd = np.zeros(len(th))
        for j in fd_idx:
            fd[j] = adj.central_fd(prob, th, which=[j])[j]
            log(f"fd[{j}] = {fd[j]:.6e}, adjoint {g[j]:.6e}")
    if a.members:
        # the batched gradient: every member (theta perturbed by up to 20%)
        # in one pdp3_adjoint_run_batch call, members concurrent on their own streams
        rng = np.random.default_rng(7)
        thetas = [th * torch.tensor(rng.uniform(0.8, 1.2, len(th))) for _ in range(a.members)]
        t3 = time.perf_counter()
        phis, grads = prob.batch_gradients(thetas, capi.PREC_FP32)
        t4 = time.perf_counter()
        log(f"batched gradient of {a.members} members: {t4 - t3:.2f} s, phi range {phis.min():.4e} .. {phis.max():.4e}")
    out = {
        "metric": "adjoint gradient of the axis E_z^2 objective (17 parameters)",
        "value": (t2 - t1), "unit": "s per gradient", "higher_is_better": False, "n_gpus": 1,
        "config": {"workload": "cfg5: 128x1024 cells, 1500 steps, 20 ppc ramp, 16-bump bunch, "
                               "FP32 forward, checkpoints every 50 steps", "particles":
                   spec.ppc * spec.cells_r * spec.cells_z + spec.bunch_macro},
        "forward_only_s": t1 - t0, "phi": float(phi), "phi_forward_only": phi_fwd,
        "gradient": g.tolist(),
        "error_fp32_vs_fp64": err,
    }
    if fd is not None:
        out["fd_fp64"] = {str(j): fd[j] for j in fd_idx}
        out["fd_rel_err"] = {str(j): abs(g[j] - fd[j]) / abs(fd[j]) for j in fd_idx}
    if a.members:
        out["members"] = a.members
        out["members_batch_s"] = t4 - t3
        out["members_s_per_gradient"] = (t4 - t3) / a.members
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: one sort interval, 50; cfg4: 80)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="cfg3")
    ap.add_argument("--sort-interval", type=int, default=None,
                    help="cell sort of the plasma every N steps (counted in the timed steps); "
                         "default 50, 80 for the cfg4 ensemble")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 plain-form comparison run")
    ap.add_argument("--error-steps", type=int, default=None,
                    help="steps of the FP32-vs-FP64 error leg (default: the configuration's own count, "
                         "1000 for cfg3, 2000 for cfg2; 0 skips it)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (counter calibration)")
    ap.add_argument("--instances", type=int, default=128, help="cfg4: ensemble size")
    ap.add_argument("--loaded-bunches", action="store_true",
                    help="cfg4: bunches loaded inside the grid at t = 0 instead of injected over time")
    ap.add_argument("--fd", default="", help="cfg5: comma-separated parameter indices to check by differences")
    ap.add_argument("--members", type=int, default=0, help="cfg5: gradients of this many perturbed members")
    ap.add_argument("--adj-steps", type=int, default=0, help="cfg5: override the step count (default 1500)")
    a = ap.parse_args()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(ensemble instances and z-slabs of one grid would shard across GPUs)", file=sys.stderr)
        sys.exit(2)

    # one sort per timed interval; the intervals are the measured optima of
    # sort cost against the particle kernel's slowdown as particles leave
    # their cells' order (flat within a few per mille over 50-100 steps at
    # cfg3, 80 for the ensemble)
    if a.sort_interval is None:
        a.sort_interval = 80 if a.config == "cfg4" else 50
    if a.steps is None:
        a.steps = a.sort_interval
    if a.config == "cfg4":
        bench_ensemble(a)
        return
    if a.config == "cfg5":
        bench_adjoint(a)
        return

    from pdp3_amd import capi
    from pdp3_amd import workloads as wl

    w = wl.config(a.config)
    n_plasma = sum(w.plasma_count(p) for p in w.plasma)
    cells = w.cells_r * w.cells_z
    print(f"# workload {w.name}: {w.cells_r}x{w.cells_z} cells, {n_plasma} plasma particles, "
          f"{sum(len(b.r) for b in w.bunches)} bunch particles, steps {a.steps}, warmup {a.warmup}",
          file=sys.stderr, flush=True)

    r32, ez32 = run_form(wl, capi, w, capi.PREC_FP32, a.steps, a.warmup, a.sort_interval)
    print(f"# fp32: {r32}", file=sys.stderr, flush=True)
    value = n_plasma * a.steps / r32["wall_s"]
    ms_per_step = 1e3 * r32["wall_s"] / a.steps

    fp64 = None
    err = None
    if not a.no_fp64:
        r64, ez64 = run_form(wl, capi, w, capi.PREC_FP64, a.steps, a.warmup, a.sort_interval)
        print(f"# fp64: {r64}", file=sys.stderr, flush=True)
        v64 = n_plasma * a.steps / r64["wall_s"]
        den = np.sqrt(np.mean(ez64 ** 2))
        err = dict(E_z_rel_rms=float(np.sqrt(np.mean((ez32 - ez64) ** 2)) / den),
                   E_z_rel_max=float(np.max(np.abs(ez32 - ez64)) / np.max(np.abs(ez64))),
                   after_steps=a.steps + a.warmup + AUTOTUNE_STEPS,
                   reference="FP64 plain form on the GPU (oracle-parity checked)")
        err_steps = w.steps if a.error_steps is None else a.error_steps
        if err_steps > 0:
            # at the configuration's own step count: E_z on the full grid, and
            # momenta and E_z on a z-slab (no sort, particles pair by index)
            err["at_config_steps"] = ez_error_full(wl, capi, w, err_steps, a.sort_interval)
            print(f"# E_z error after {err_steps} steps: {err['at_config_steps']}", file=sys.stderr, flush=True)
        err["particles"] = momentum_error(wl, capi, a.config, min(1024, w.cells_z), max(err_steps, a.steps + a.warmup))
        print(f"# particle errors: {err['particles']}", file=sys.stderr, flush=True)
        p64 = r64["particle_ms"] / max(r64["particle_launches"], 1)
        n_sp = max(len(w.plasma), 1)
        fp64 = dict(value=v64, ms_per_step=1e3 * r64["wall_s"] / a.steps, particle_kernel_ms=p64,
                    roofline_frac=(n_plasma * BYTES_PER_PARTICLE["fp64"] + n_sp * cells * BYTES_PER_CELL) / n_sp
                    / (p64 * 1e-3) / 1e9 / HBM_PEAK_GBS)

    # one launch per plasma species and step: the average launch moves its
    # species' particles and the grid once
    n_species = max(len(w.plasma), 1)
    per_launch_ms = r32["particle_ms"] / max(r32["particle_launches"], 1)
    bytes_launch = (n_plasma * BYTES_PER_PARTICLE["fp32"] + n_species * cells * BYTES_PER_CELL) / n_species
    achieved = bytes_launch / (per_launch_ms * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                    traffic=traffic_from_profile(f"k_fused32_tiled_{a.config}"), kernel="k_fused32_tiled (plasma species)",
                    bytes_per_launch=bytes_launch, avg_launch_ms=per_launch_ms)

    cpu = None
    if not a.no_cpu:
        threads, visible, model = host_cpu()
        cpu = cpu_baseline(wl, capi, w, threads, visible, model)
        print(f"# cpu: {cpu}", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": describe_workload(w),
                   "precision_form": "fp32 cell-relative (FP64 grid sums, axis cell, fields)",
                   "sort_interval": a.sort_interval, "tile_ppt_waves": r32["tile"]},
        "fp64_plain": fp64,
        "speedup_fp32_vs_fp64": (value / fp64["value"]) if fp64 else None,
        "error_fp32_vs_fp64": err,
        "roofline": roofline,
        # device time per timed step by kernel group (HIP events on the
        # library's stream); the rest of ms_per_step is resets and launch gaps
        "device_ms_per_step": {k: r32[k + "_ms"] / a.steps for k in ("particle", "bunch", "h", "e", "sort")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

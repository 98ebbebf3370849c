"""Benchmark of the fused D3Q27 periodic-shift step on one MI355X.

Workload (the sphere configuration): 256 x 128 x 128 channel, icosphere of
diameter 32 cells (5120 triangles) voxelized through the STL path, velocity
inlet U = 0.04 (moment-based), pressure outlet, periodic y (copies) and z
(implicit in the shift), Re = 1000 (nu = 1.28e-3), Smagorinsky Cs = 0.17.
A step is one updateGrid pass: applyStreaming, applyBCPeriodicY,
applyLocalCellUpdate (bulk + boundary kernels).  Inputs are resident in HBM
before timing starts.

value = MLUPS of the FP32-storage step (cellCount * steps / wall time between
device synchronisations).  FP16 / BF16 deviation-storage variants and the
64-domain batch are reported under "variants".  roofline.achieved =
algorithmic bytes of the dominant launch (216 B FP32 / 108 B 16-bit per
updated cell x the cells it updates: bulk cells for k_ps_bulk, bulk + listed
boundary cells for k_ps_fused) divided by its mean duration from HIP events
on its stream around each of its launches inside as many whole steps as
the timed run (repeated after it, so that the events stay out of its clock).

cpu_baseline: the project's FP32 CPU oracle (OpenMP) on the same workload for
a bounded number of steps on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "MLUPS + % HBM roofline, D3Q27 fused step: plain FP32 vs FP16-store/batched/MLP"
HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = {"f32": 216, "f16": 108, "bf16": 108}
INFINITY_CACHE_BYTES = 256 << 20  # MI355X last-level (MALL) cache


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_SUMMARY = os.path.join(REPO, "profiles", "sphere_256x128x128_hbm_counters.md")


def measured_traffic(kernel: str, storage: str):
    """HBM bytes per launch of `kernel` at `storage` from the committed
    rocprofv3 counter summary (FETCH_SIZE x 2 + WRITE_SIZE x 1, the factors
    from a copy kernel of known size); None when not measured."""
    import re

    try:
        text = open(TRAFFIC_SUMMARY).read()
    except OSError:
        return None
    pat = re.compile(rf"^traffic {re.escape(kernel)} {re.escape(storage)} = (\d+)", re.M)
    m = pat.search(text)
    return int(m.group(1)) if m else None


def build_sphere_grid(P, cases, storage: str, shape=(256, 128, 128), u_inlet=0.04, reynolds=1000.0, cs=0.17):
    nx, ny, nz = shape
    nu = cases.sphere_nu(u_inlet, cases.SPHERE_DIAMETER, reynolds)
    info = P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=nu, res=1.0)
    tri = cases.icosphere(4, cases.SPHERE_DIAMETER / 2, cases.SPHERE_CENTER)
    stl = P.STLStruct.from_triangles(tri)
    mask = P.BoolArray(nx * ny * nz)
    P.applyMarkersInsideSTL(mask, stl, True, info)
    solid = mask.get().astype(bool)
    flags = cases.channel_flags(shape, solid)
    g = P.GridStruct(info, storage)
    g.set_flags(flags)
    g.set_given_values(cases.channel_given_values(u_inlet))
    g.set_smagorinsky(cs)
    g.set_periodic_copies(y=True)
    P.fill_equilibrium(g)
    return g, flags, int(solid.sum()), nu


def time_steps(P, g, steps: int, warmup: int):
    P.updateGrid([g], warmup)
    P.synchronize()
    t0 = time.perf_counter()
    P.updateGrid([g], steps)
    P.synchronize()
    return time.perf_counter() - t0


def in_step_kernel_ms(P, g, steps: int) -> float:
    """Mean duration of the grid's dominant update kernel inside `steps` whole
    steps (streaming, periodic copies, update): an event pair on the grid's
    stream around each of its launches.  Run after the timed steps, so that
    the events do not touch the timed region's wall clock."""
    g.profile(True)
    P.updateGrid([g], steps)
    total_ms, launches = g.profile_read()
    g.profile(False)
    return total_ms / launches


def batch_leg(P, cases, steps: int, autotune_steps: int, storage: str = "f16", avg_steps: int = 2500,
              avg_every: int = 100):
    """64 sphere-in-channel domains of uneven size advanced by replays of a
    hipGraph of 1000 captured steps; aggregate MLUPS, its HBM fraction, and
    per-domain drag coefficient averaged over avg_steps steps after the timed
    run.  The autotuner searches kernel variant, grid order, cells per thread
    and tile order; the multi-step graph is captured in an untimed warm-up."""
    doms = cases.batch_domains()
    grids = [cases.domain_grid(d, storage)[0] for d in doms]
    b = P.Batch(grids)
    tune = None
    if autotune_steps:
        (t, o, cpt, tile), times = b.autotune_layouts(autotune_steps)
        tune = {"tuning": t, "order": o, "cells_per_thread": cpt, "tile_order": tile,
                "ms_per_step": {f"tiles{ti}_cpt{c + 1}_order{od}": [round(float(x), 4) for x in times[ti, c, od]]
                                for ti in range(2) for c in range(2) for od in range(2)}}
    cells, blocks, nbnd = b.census()
    if steps >= b.steps_per_graph:
        b.step(b.steps_per_graph)  # captures the multi-step graph
    for g in grids:  # restart from rest after tuning and warm-up
        P.fill_equilibrium(g)
    P.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    P.synchronize()
    el = time.perf_counter() - t0
    # drag: after the timed steps, sampled every avg_every steps over avg_steps
    # more (untimed), averaged per domain
    boxes = [cases.sphere_box_of(d["center"], d["diameter"]) for d in doms]
    samples = []
    for _ in range(avg_steps // avg_every):
        b.step(avg_every)
        samples.append([g.momentum_exchange(lo, hi)[0] for g, (lo, hi) in zip(grids, boxes)])
    fx = np.mean(np.array(samples), axis=0) if samples else np.zeros(len(doms))
    cd = [cases.drag_coefficient(f, d["u_inlet"], d["diameter"]) for f, d in zip(fx, doms)]
    b.close()
    for g in grids:
        g.close()
    mlups = cells * steps / el / 1e6
    achieved = mlups * 1e6 * BYTES_PER_UPDATE[storage] / 1e9
    return {"mlups": mlups, "ms_per_step": el * 1e3 / steps, "domains": len(doms),
            "cells": cells, "boundary_cells": nbnd, "steps": steps, "storage": storage, "autotune": tune,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "bytes_per_cell": BYTES_PER_UPDATE[storage],
                         "note": "whole step (all kernels of the replayed graph) over all domains"},
            "working_set_bytes": cells * 27 * (4 if storage == "f32" else 2),
            "drag_coefficient_mean": [round(float(c), 5) for c in cd],
            "drag_average": {"after_step": steps, "steps": avg_steps, "every": avg_every},
            "L": [d["L"] for d in doms], "re": [round(d["re"], 1) for d in doms]}


def drag_history_leg(P, cases, storages, steps: int, every: int, tuning=None):
    """The sphere configuration's drag history (flowAroundSphereMainSingle.cu:
    240-252 records C_D every step): `steps` steps from rest per storage, the
    momentum-exchange drag coefficient every `every` steps, and its mean over
    the second half of the run.  The history in the output is thinned to every
    200 steps."""
    out = {}
    lo, hi = cases.sphere_box()
    for st in storages:
        g, _, _, _ = build_sphere_grid(P, cases, st)
        if tuning and st in tuning:
            g.tuning = tuning[st]
        fx = []
        t0 = time.perf_counter()
        for _ in range(steps // every):
            P.updateGrid([g], every)
            fx.append(float(g.momentum_exchange(lo, hi)[0]))
        el = time.perf_counter() - t0
        g.close()
        cd = np.array([cases.drag_coefficient(f) for f in fx])
        thin = max(1, 200 // every)
        out[st] = {"steps": steps, "every": every, "seconds": el,
                   "mean_cd_second_half": float(np.mean(cd[len(cd) // 2:])),
                   "std_cd_second_half": float(np.std(cd[len(cd) // 2:])),
                   "cd_every_200_steps": [round(float(c), 5) for c in cd[thin - 1::thin]]}
    if "f32" in out:
        for st in storages:
            if st != "f32":
                out[st]["mean_cd_rel_to_f32"] = out[st]["mean_cd_second_half"] / out["f32"]["mean_cd_second_half"] - 1
    out["note"] = ("D = 32 cells; the reference publishes C_D(Re=1000) = 0.514 at D/res = 160 "
                   "(flowAroundSphere/resultsAllPlotter.py:188-191), a finer grid")
    return out


def access_probes(P, n: int = 256):
    """The periodic-shift access pattern taken apart on an n^3 grid (no
    collision): stream count, reads against writes, line alignment of the
    streams, cells per lane.  GB/s of bytes moved per launch."""
    rows = []

    def run(storage, streams, mode, cpl, aligned):
        gbs = P.probe_ps_access(n, n, n, storage, streams, mode, cpl, aligned)
        rows.append({"storage": storage, "streams": streams, "mode": mode, "cells_per_lane": cpl,
                     "line_aligned": aligned, "GBs": gbs})
        log(f"  {storage} {mode:>7} streams {streams:2d} cells/lane {cpl} {'aligned' if aligned else 'shifted':>7}: "
            f"{gbs:7.0f} GB/s")

    for storage in ("f32", "f16"):
        for streams in (1, 2, 4, 8, 16, 27):
            run(storage, streams, "inplace", 1, False)
        for mode in ("read", "write"):
            for aligned in (False, True):
                run(storage, 27, mode, 1, aligned)
        run(storage, 27, "inplace", 1, True)
        for cpl in (2, 4):
            if storage == "f32":
                run(storage, 27, "inplace", cpl, False)
            run(storage, 27, "inplace", cpl, True)
    return rows


def occupancy_sweep(P, cases, n: int, steps: int = 200, repeats: int = 3):
    """Bulk workgroups per CU capped at 6 to 2 (0: uncapped) on the sphere
    256x128x128 and the periodic n^3 box, FP32 and FP16, with the variant the
    autotuner picks uncapped: the bulk kernel alone at the shift reached (50
    launches back to back) and whole steps, best of `repeats` passes over the
    caps in alternating order."""
    rows = []
    caps = (0, 6, 5, 4, 3, 2)
    for name in ("sphere", "box"):
        for storage in ("f32", "f16"):
            if name == "sphere":
                g = build_sphere_grid(P, cases, storage)[0]
            else:
                nx, ny, nz = cases.tgv_shape(n)
                g = P.GridStruct(P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=0.005), storage)
                g.set_periodic_copies(x=True, y=True)
                P.fill_equilibrium(g, *cases.tgv_fields(n))
            g.autotune(5)
            g.occupancy = 0
            best = {c: [float("inf"), float("inf")] for c in caps}
            for r in range(repeats):
                for c in (caps if r % 2 == 0 else caps[::-1]):
                    g.occupancy = c
                    ms_step = time_steps(P, g, steps, 20) * 1e3 / steps
                    ms_kernel = g.time_update_kernel(50)[0]
                    best[c][0] = min(best[c][0], ms_kernel)
                    best[c][1] = min(best[c][1], ms_step)
            for c in caps:
                k, s = best[c]
                rows.append({"grid": name, "storage": storage, "cells": g.n, "tuning": g.tuning,
                             "workgroups_per_cu": c, "bulk_kernel_us": k * 1e3, "step_us": s * 1e3,
                             "kernel_vs_uncapped": best[0][0] / k, "step_vs_uncapped": best[0][1] / s})
                log(f"  {name} {storage} variant {g.tuning} cap {c or 'none':>4}: bulk kernel {k * 1e3:7.1f} us "
                    f"({best[0][0] / k:5.3f}x), step {s * 1e3:7.1f} us ({best[0][1] / s:5.3f}x)")
            g.close()
    return rows


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def copy_probe(P, mib: int = 2048, reps: int = 20):
    """A plain device copy of `mib` MiB (beyond the 256 MB Infinity Cache): the
    HBM rate a streaming kernel reaches on this device, read + write bytes."""
    nbytes = mib << 20
    a, b = P.DeviceBuffer(nbytes), P.DeviceBuffer(nbytes)
    P.device_copy(b, a, nbytes)
    P.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        P.device_copy(b, a, nbytes)
    P.synchronize()
    el = time.perf_counter() - t0
    a.close()
    b.close()
    return {"bytes_per_copy": 2 * nbytes, "GBs": 2 * nbytes * reps / el / 1e9, "copies": reps}


def large_box_leg(P, cases, n: int, steps: int):
    """The bulk kernel on a fully periodic n^3 box ((n+1)^2 n cells; >= 1 GB of
    FP32 distributions at n = 256, beyond the 256 MB Infinity Cache), Taylor-Green
    initial field, no Smagorinsky: MLUPS and the bulk kernel's HBM rate."""
    out = {}
    nx, ny, nz = cases.tgv_shape(n)
    rho, ux, uy, uz = cases.tgv_fields(n)
    for storage in ("f32", "f16"):
        info = P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=0.005)
        g = P.GridStruct(info, storage)
        g.set_periodic_copies(x=True, y=True)
        P.fill_equilibrium(g, rho, ux, uy, uz)
        g.autotune(5)  # variant and occupancy cap (the decaying flow advances a few steps)
        el = time_steps(P, g, steps, 10)
        ms = in_step_kernel_ms(P, g, 256)
        ms_b2b, _ = g.time_update_kernel(256, streaming=True)
        gbs = BYTES_PER_UPDATE[storage] * g.n / (ms * 1e-3) / 1e9
        out[storage] = {"cells": g.n, "f_bytes": g.n * 27 * (4 if storage == "f32" else 2),
                        "mlups": g.n * steps / el / 1e6, "bulk_kernel_ms": ms, "achieved_GBs": gbs,
                        "bulk_kernel_ms_back_to_back": ms_b2b,
                        "roofline_frac": gbs / HBM_PEAK_GBS, "tuning": g.tuning,
                        "occupancy": g.occupancy}
        g.close()
        # the pattern's ceiling under the cap the bulk kernel ran with
        out[storage]["pattern_limit_GBs"] = P.probe_ps_pattern(nx, ny, nz, storage,
                                                               occupancy=out[storage]["occupancy"])
        out[storage]["frac_of_pattern_limit"] = gbs / out[storage]["pattern_limit_GBs"]
    out["f16_over_f32"] = out["f16"]["mlups"] / out["f32"]["mlups"]
    out["note"] = "working sets of 1.8 GB (FP32) and 0.9 GB (FP16): HBM-resident, beyond the 256 MB Infinity Cache"
    return out


def esotwist_leg(P, n: int, steps: int):
    """The esotwist update (DIAD grids, include/applyLocalCellUpdate.h:77-171) on a
    dense, fully periodic n^3 box: every cell stored with its seven neighbour
    indices (periodicity through the indices), uniform flow, no Smagorinsky.
    Algorithmic bytes per cell: 27 reads + 27 writes of the storage (216 / 108
    B); the seven 4-byte neighbour indices are reported as overhead."""
    out = {}
    ijk, nbr = P.DIADGridStruct.dense_box(n, n, n)
    for storage in ("f32", "f16"):
        info = P.InfoStruct(cellCountX=n, cellCountY=n, cellCountZ=n, nu=0.005)
        g = P.DIADGridStruct(info, ijk, nbr, storage)
        g.fill_equilibrium(1.0, 0.02, 0.0, 0.0)
        for _ in range(5):
            P.applyStreaming(g)
            P.applyLocalCellUpdate(g)
        P.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            P.applyStreaming(g)
            P.applyLocalCellUpdate(g)
        P.synchronize()
        el = time.perf_counter() - t0
        per_cell = BYTES_PER_UPDATE[storage]
        gbs = per_cell * g.n * steps / el / 1e9
        out[storage] = {"cells": g.n, "mlups": g.n * steps / el / 1e6, "ms_per_step": el * 1e3 / steps,
                        "bytes_per_cell": per_cell, "achieved_GBs": gbs, "roofline_frac": gbs / HBM_PEAK_GBS,
                        "overhead_bytes_per_cell": 28,
                        "achieved_GBs_with_overhead": (per_cell + 28) * g.n * steps / el / 1e9}
        g.close()
    return out


def two_level_leg(P, cases, steps: int, storage: str = "f16", autotune: bool = True):
    """The sphere at Re = 1e5 on two levels (coarse 256x128x128 + fine 256x192x192
    around the sphere); MLUPS counts the updated cells of each level times its
    steps per coarse step (sum_l N_l 2^l without the skipped covered coarse
    cells and fine ghost cells)."""
    (gc, gf), (lo, hi) = cases.two_level_sphere(storage)
    tuning = {}
    if autotune:
        # each level's local-update variant, then both restart from rest
        for name, g in (("coarse", gc), ("fine", gf)):
            tuning[name] = g.autotune(10)[0]
            P.fill_equilibrium(g)
    P.updateGrid([gc, gf], 5)
    P.synchronize()
    t0 = time.perf_counter()
    P.updateGrid([gc, gf], steps)
    P.synchronize()
    el = time.perf_counter() - t0
    bc, nc, skipped = gc.census()
    bf, nf, sf = gf.census()
    updates = (bc + nc) + 2 * (bf + nf)  # cells actually updated (covered and ghost cells are skipped)
    force = gf.momentum_exchange(lo, hi)
    out = {"mlups": updates * steps / el / 1e6, "ms_per_coarse_step": el * 1e3 / steps, "coarse_steps": steps,
           "coarse_cells": gc.n, "covered_coarse_cells": int(skipped), "fine_cells": gf.n,
           "fine_ghost_cells": int(sf), "storage": storage, "tuning": tuning,
           "fine_force_after_run": [float(x) for x in force]}
    gc.close()
    gf.close()
    return out


def diad_leg(P, cases, levels: int, steps: int, storage: str = "f16"):
    """The wall-refined DIAD sphere (D = 2000 mm on a 220^3 base grid of 100 mm,
    `levels` levels, wall refinement span 3, Re = 1e6, Cs = 0.1): the device
    builder's time and cell counts, then `steps` coarse steps of the level
    recursion.  MLUPS counts sum_l N_l 2^l cell updates per coarse step; the
    algorithmic bytes per update are the storage's 54 values (216 / 108 B); the
    seven 4-byte neighbour indices and the flag byte (29 B) are overhead,
    reported apart, and the interface transfers are in the step's time but not
    in its bytes."""
    t0 = time.perf_counter()
    lv, grids = cases.diad_sphere_reference(P, storage, levels=levels)
    setup_s = time.perf_counter() - t0
    counts = [g.n for g in grids]
    updates = sum(c << l for l, c in enumerate(counts))
    P.updateGrid(grids, 2)
    P.synchronize()
    t0 = time.perf_counter()
    P.updateGrid(grids, steps)
    P.synchronize()
    el = time.perf_counter() - t0
    per_update = BYTES_PER_UPDATE[storage]
    mlups = updates * steps / el / 1e6
    force = grids[-1].momentum_exchange()
    out = {"mlups": mlups, "ms_per_coarse_step": el * 1e3 / steps, "coarse_steps": steps, "storage": storage,
           "levels": levels, "cells_per_level": counts, "updates_per_coarse_step": updates,
           "build_ms": lv.build_ms, "setup_s": setup_s, "bytes_per_update": per_update,
           "achieved_GBs": mlups * 1e6 * per_update / 1e9,
           "roofline_frac": mlups * 1e6 * per_update / 1e9 / HBM_PEAK_GBS,
           "overhead_bytes_per_update": 29,
           "achieved_GBs_with_overhead": mlups * 1e6 * (per_update + 29) / 1e9,
           "finest_force_after_run": [float(x) for x in force]}
    for g in grids:
        g.close()
    lv.close()
    return out


def scalar_leg(P, cases, steps: int):
    """Scalar transport on the sphere channel (256x128x128, FP32), the scalar
    driver's step: streaming, the zero-gradient scalar outlet on the last x
    layer, then f and the 27 scalar populations updated in one launch; the
    algorithmic bytes are 2 x 216 B per cell update (the outlet's layer copy is
    left out, 0.4 % of it)."""
    g, _, _, _ = build_sphere_grid(P, cases, "f32")
    st = P.ScalarTransportStruct(g, 0.6)
    given = np.full(g.n, np.nan, np.float32)
    given[np.arange(g.n) % 256 == 0] = 1.0  # scalar prescribed at the inlet
    st.set_given(given)
    for _ in range(5):
        P.applyStreaming(g)
        P.applyZeroGradientOutletScalarTransportBC(g, st)
        P.applyScalarTransportLocalCellUpdate(g, st)
    P.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        P.applyStreaming(g)
        P.applyZeroGradientOutletScalarTransportBC(g, st)
        P.applyScalarTransportLocalCellUpdate(g, st)
    P.synchronize()
    el = time.perf_counter() - t0
    mlups = g.n * steps / el / 1e6
    out = {"mlups": mlups, "ms_per_step": el * 1e3 / steps, "bytes_per_update": 432,
           "achieved_GBs": mlups * 1e6 * 432 / 1e9, "roofline_frac": mlups * 1e6 * 432 / 1e9 / HBM_PEAK_GBS}
    g.close()
    return out


def closure_leg(P, n: int, log, closure_every: int = 8, sphere_steps: int = 20000):
    """Learned closure at the decaying-turbulence configuration: n^3 fine run
    (FP16 storage, Cs 0.1), filtered 8x to (n/8)^3 every 200 steps (a coarse
    grid on which an SGS model matters: without one the coarse spectrum error
    is 45 % higher), network trained on the filtered pairs, coarse LES runs
    judged by energy spectrum against the filtered fine run (the network
    evaluated every step and every `closure_every` steps), and the
    bulk-kernel cost with the network against Smagorinsky."""
    from periodicshiftlbm_amd import les

    wf = les.Workflow(P, n=n, r=8, k_peak=n / 64.0, nu_fine=5e-4, cs_fine=0.1, storage="f16", seed=7, log=log)
    steps = 3200 if n >= 256 else 800
    every = 200 if n >= 256 else 100
    wf.run_fine(steps=steps, every=every, first_sample=every)
    wf.train(iters=2000, batch=16384, lr=2e-3)
    out = {"fine": f"{n}^3", "coarse": f"{n // 8}^3", "fine_steps": steps, "snapshot_every": every,
           "fine_mlups": wf.timings.get("fine_mlups"), "train_s": wf.timings.get("train_s"),
           "targets": wf.target_stats, "loss_first": float(wf.loss_hist[0]), "loss_last": float(wf.loss_hist[-1])}
    for k in sorted({1, closure_every}):
        wf.closure_every = k
        res = wf.compare(start=every)
        r32, t32 = wf.cost_ratio(n=256, storage="f32", cs=res["smagorinsky"]["param"])
        r16, t16 = wf.cost_ratio(n=256, storage="f16", cs=res["smagorinsky"]["param"])
        log(f"  closure (network every {k} steps): resolved-band spectrum error none "
            f"{res['none']['evaluation_resolved']:.4f}, smagorinsky(Cs {res['smagorinsky']['param']}) "
            f"{res['smagorinsky']['evaluation_resolved']:.4f}, mlp {res['mlp']['evaluation_resolved']:.4f}, tuned "
            f"smagorinsky {res['smagorinsky_tuned']['evaluation_resolved']:.4f}; gain vs smagorinsky "
            f"{res['gain_vs_smagorinsky'] * 100:.1f} %, vs none {res['gain_vs_none'] * 100:.1f} %, vs tuned "
            f"smagorinsky {res['gain_vs_smagorinsky_tuned'] * 100:.1f} %; cost ratio f32 {r32:.2f} f16 {r16:.2f}")
        out[f"every_{k}"] = {"spectrum_error": res, "gain_vs_smagorinsky": res["gain_vs_smagorinsky"],
                             "gain_vs_none": res["gain_vs_none"],
                             "gain_vs_smagorinsky_tuned": res["gain_vs_smagorinsky_tuned"],
                             "cost_ratio_vs_smagorinsky": {"f32": r32, "f16": r16},
                             "bulk_kernel_ms_256cubed": {"f32": t32, "f16": t16}}
    head = out[f"every_{closure_every}"]
    out.update(closure_every=closure_every, gain_vs_smagorinsky=head["gain_vs_smagorinsky"],
               gain_vs_none=head["gain_vs_none"], cost_ratio_vs_smagorinsky=head["cost_ratio_vs_smagorinsky"])
    if sphere_steps:
        # the same network on the Re = 1e5 sphere against a 2x finer Smagorinsky run
        sph = les.sphere_comparison(P, wf.net, coarse_steps=sphere_steps, closure_every=closure_every, log=log)
        out["sphere_re1e5"] = sph
        parts = []
        for name in ("none", "smagorinsky", "mlp"):
            r = sph[name]
            parts.append(f"{name} " + (f"diverged at step {r['diverged_at']}" if "diverged_at" in r else
                                      f"C_D {r['cd_mean']:.3f} ({r['cd_rel_err'] * 100:.1f} %), wake u_x "
                                      f"{r['wake_ux_rel_l2'] * 100:.1f} %"))
        log(f"  sphere Re 1e5 against the 2x finer run (C_D {sph['fine'].get('cd_mean', float('nan')):.3f}): "
            + "; ".join(parts))
    return out


def cpu_baseline(cases, flags, nu, cs, shape, budget_s: float):
    import oracle as O

    nx, ny, nz = shape
    og = O.Grid(nx, ny, nz, dtype=np.float32, nu=nu, cs=cs, flags=flags,
                bcvals=cases.channel_given_values(0.04))
    og.fill_equilibrium()
    og.step(periodic_y=True)  # warm caches / threads
    steps = 0
    t0 = time.perf_counter()
    while True:
        og.step(periodic_y=True)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 200:
            break
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": nx * ny * nz * steps / el / 1e6, "unit": "MLUPS", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(), "build": "g++ " + O.build_flags(),
            "sample": f"{steps} FP32 oracle steps of the same 256x128x128 sphere workload ({el:.1f} s, OpenMP)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--storages", default="f32,f16,bf16")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", type=int, default=None,
                    help="use this local-update variant (bits: 1 NT loads, 2 NT stores, 4 XCD order, 8 fused "
                         "boundary) instead of autotuning")
    ap.add_argument("--batch-steps", type=int, default=5000, help="steps of the 64-domain batch (0: skip)")
    ap.add_argument("--drag-steps", type=int, default=20000,
                    help="steps of the sphere drag history per storage (0: skip)")
    ap.add_argument("--two-level-steps", type=int, default=200, help="coarse steps of the two-level sphere (0: skip)")
    ap.add_argument("--large-n", type=int, default=256, help="edge of the large periodic box (0: skip)")
    ap.add_argument("--scalar-steps", type=int, default=100,
                    help="steps of the scalar-transport sphere channel (0: skip)")
    ap.add_argument("--diad-steps", type=int, default=20,
                    help="coarse steps of the wall-refined DIAD sphere (0: skip)")
    ap.add_argument("--diad-levels", type=int, default=6, help="levels of the DIAD sphere")
    ap.add_argument("--diad-storages", default="f16,f32", help="storages of the DIAD sphere leg")
    ap.add_argument("--closure-n", type=int, default=512,
                    help="fine grid edge of the learned-closure leg (0: skip)")
    ap.add_argument("--closure-sphere-steps", type=int, default=20000,
                    help="coarse steps of the Re 1e5 sphere comparison of the closure (0: skip)")
    ap.add_argument("--closure-every", type=int, default=8,
                    help="network evaluated every this many steps in the learned-closure leg")
    ap.add_argument("--access-probes", action="store_true",
                    help="only run the access-pattern probes on a 256^3 grid and print their JSON line")
    ap.add_argument("--occupancy-sweep", action="store_true",
                    help="only run the bulk occupancy-cap sweep (sphere and large box, FP32 and FP16) "
                         "and print its JSON line")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB (10 times) for counter calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "multi-GPU sharding is not part of it.", file=sys.stderr)
        sys.exit(2)

    import periodicshiftlbm_amd as P
    from periodicshiftlbm_amd import cases

    P.require_device()
    if args.access_probes:
        log("[bench] periodic-shift access probes, 256^3")
        cp = copy_probe(P)
        log(f"  device copy of 2 GiB: {cp['GBs']:.0f} GB/s")
        print(json.dumps({"access_probes": access_probes(P), "hbm_copy_probe": cp}), flush=True)
        return
    if args.occupancy_sweep:
        log(f"[bench] bulk occupancy caps, sphere 256x128x128 and periodic box {args.large_n}^3")
        print(json.dumps({"occupancy_sweep": occupancy_sweep(P, cases, args.large_n)}), flush=True)
        return
    if args.copy_calibration_mb:
        nbytes = args.copy_calibration_mb << 20
        a, b = P.DeviceBuffer(nbytes), P.DeviceBuffer(nbytes)
        for _ in range(10):
            P.device_copy(b, a, nbytes)
        P.synchronize()
        log(f"[bench] copied {nbytes} bytes 10 times")
        return
    shape = (256, 128, 128)
    ncell = int(np.prod(shape))
    results = {}
    flags_keep = None
    nu_keep = None
    for storage in args.storages.split(","):
        log(f"[bench] storage={storage} shape={shape}")
        g, flags, nsolid, nu = build_sphere_grid(P, cases, storage, shape)
        flags_keep, nu_keep = flags, nu
        tune_times = None
        if args.tuning is not None:
            g.tuning = args.tuning
        elif not args.no_autotune:
            best, tt = g.autotune(10)
            tune_times = {int(v): float(t) for v, t in enumerate(tt)}
            P.fill_equilibrium(g)  # restart from rest after tuning
            log(f"  autotune: variant {best}, workgroups per CU {g.occupancy or 'uncapped'} "
                f"({', '.join(f'{v}:{t * 1e3:.1f}us' for v, t in tune_times.items())})")
        el = time_steps(P, g, args.steps, args.warmup)
        mlups = ncell * args.steps / el / 1e6
        bulk, nbc, skipped = g.census()
        lo, hi = cases.sphere_box()
        fx = g.momentum_exchange(lo, hi)[0]
        # the dominant launch (k_ps_fused updates the bulk and the boundary
        # list, k_ps_bulk the bulk only) inside as many steps again, an event
        # pair around each launch; beside it, 256 launches back to back with
        # the shifter advancing (faster: no other kernels between them)
        avg_ms = in_step_kernel_ms(P, g, args.steps)
        b2b_ms, kernel = g.time_update_kernel(256, streaming=True)
        cells_per_launch = bulk + (nbc if kernel == "k_ps_fused" else 0)
        bytes_per_launch = BYTES_PER_UPDATE[storage] * cells_per_launch
        achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        pattern = P.probe_ps_pattern(*shape, storage, occupancy=g.occupancy)
        results[storage] = {
            "mlups": mlups, "ms_per_step": el * 1e3 / args.steps, "bulk_kernel_ms": avg_ms,
            "bulk_kernel_ms_back_to_back": b2b_ms,
            "kernel": kernel, "cells_per_launch": cells_per_launch,
            "pattern_limit_GBs": pattern, "frac_of_pattern_limit": achieved / pattern,
            "bulk_cells": bulk, "boundary_cells": nbc, "solid_cells": nsolid,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS,
                         "traffic": measured_traffic(kernel, storage)},
            "drag_force_x_after_run": fx,
            "tuning": g.tuning, "occupancy": g.occupancy, "autotune_ms": tune_times,
            "working_set_bytes": ncell * (27 * (4 if storage == "f32" else 2) + 1),
        }
        log(f"  {mlups:.1f} MLUPS, {kernel} {avg_ms * 1e3:.1f} us ({achieved:.0f} GB/s, "
            f"{achieved / HBM_PEAK_GBS * 100:.1f}% of HBM peak, {achieved / pattern * 100:.0f}% of the access "
            f"pattern's {pattern:.0f} GB/s)")
        g.close()

    head = results.get("f32") or next(iter(results.values()))
    head_storage = "f32" if "f32" in results else next(iter(results))
    out = {
        "metric": METRIC,
        "value": head["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head_storage,
        "data": "synthetic",
        "config": {"workload": "sphere channel 256x128x128, icosphere D=32 (5120 triangles) voxelized by the STL "
                               "path, velocity inlet U=0.04, pressure outlet, periodic y/z, Re=1000, Cs=0.17; "
                               "step = streaming + periodic-y copy + fused bulk + boundary kernels",
                   "cells": ncell, "storage": head_storage},
        "roofline": dict(head["roofline"], kernel=head["kernel"], bytes_per_cell=BYTES_PER_UPDATE[head_storage],
                         cells_per_launch=head["cells_per_launch"], avg_launch_ms=head["bulk_kernel_ms"],
                         avg_launch_ms_back_to_back=head["bulk_kernel_ms_back_to_back"]),
        "variants": {k: {"mlups": v["mlups"], "ms_per_step": v["ms_per_step"], "roofline_frac": v["roofline"]["frac"],
                         "achieved_GBs": v["roofline"]["achieved"], "kernel": v["kernel"],
                         "pattern_limit_GBs": v["pattern_limit_GBs"], "frac_of_pattern_limit": v["frac_of_pattern_limit"],
                         "kernel_ms": v["bulk_kernel_ms"], "kernel_ms_back_to_back": v["bulk_kernel_ms_back_to_back"],
                         "tuning": v["tuning"], "occupancy": v["occupancy"], "autotune_ms": v["autotune_ms"],
                         "working_set_bytes": v["working_set_bytes"],
                         "cache_resident": v["working_set_bytes"] < INFINITY_CACHE_BYTES}
                     for k, v in results.items()},
    }
    if "f32" in results and "f16" in results:
        out["variants"]["f16_over_f32"] = results["f16"]["mlups"] / results["f32"]["mlups"]
        out["variants"]["f16_over_f32_note"] = ("sphere: the FP16 working set fits the 256 MB Infinity Cache; "
                                                "large_box.f16_over_f32 is the HBM-resident ratio")
    if args.drag_steps:
        log(f"[bench] sphere drag history, {args.drag_steps} steps per storage")
        dh = drag_history_leg(P, cases, args.storages.split(","), args.drag_steps, 20,
                              {k: v["tuning"] for k, v in results.items()})
        log("  " + ", ".join(f"{k}: mean C_D {v['mean_cd_second_half']:.4f} ({v['seconds']:.1f} s)"
                             for k, v in dh.items() if isinstance(v, dict)))
        out["variants"]["drag_history"] = dh
    if args.batch_steps:
        log(f"[bench] batch of 64 domains, {args.batch_steps} steps")
        bl = batch_leg(P, cases, args.batch_steps, 0 if (args.no_autotune or args.tuning is not None) else 3)
        log(f"  {bl['mlups']:.1f} MLUPS over {bl['cells']} cells")
        out["variants"]["batch"] = bl
    if args.large_n:
        cp = copy_probe(P)
        log(f"[bench] device copy of 2 GiB: {cp['GBs']:.0f} GB/s (read + write)")
        out["variants"]["hbm_copy_probe"] = cp
        log(f"[bench] large periodic box {args.large_n}^3")
        lb = large_box_leg(P, cases, args.large_n, 100)
        log("  " + ", ".join(f"{k}: {v['mlups']:.0f} MLUPS, {v['achieved_GBs']:.0f} GB/s "
                             f"({v['frac_of_pattern_limit'] * 100:.0f} % of the access pattern's "
                             f"{v['pattern_limit_GBs']:.0f} GB/s)" for k, v in lb.items() if isinstance(v, dict))
            + f"; FP16 / FP32 {lb['f16_over_f32']:.2f}")
        out["variants"]["large_box"] = lb
        log(f"[bench] esotwist (DIAD) dense periodic box {args.large_n}^3")
        eso = esotwist_leg(P, args.large_n, 50)
        log("  " + ", ".join(f"{k}: {v['mlups']:.0f} MLUPS, {v['achieved_GBs']:.0f} GB/s" for k, v in eso.items()))
        out["variants"]["esotwist"] = eso
    if args.two_level_steps:
        log(f"[bench] two-level sphere, {args.two_level_steps} coarse steps")
        tl = two_level_leg(P, cases, args.two_level_steps, autotune=not args.no_autotune)
        log(f"  {tl['mlups']:.1f} MLUPS ({tl['ms_per_coarse_step']:.3f} ms per coarse step, variants {tl['tuning']})")
        out["variants"]["two_level"] = tl
    if args.scalar_steps:
        log(f"[bench] scalar transport, sphere channel, {args.scalar_steps} steps")
        sl = scalar_leg(P, cases, args.scalar_steps)
        log(f"  {sl['mlups']:.0f} MLUPS, {sl['achieved_GBs']:.0f} GB/s (f and scalar, 432 B per cell)")
        out["variants"]["scalar_transport"] = sl
    if args.diad_steps:
        log(f"[bench] wall-refined DIAD sphere, {args.diad_levels} levels, {args.diad_steps} coarse steps")
        dl = {}
        for st in args.diad_storages.split(","):
            dl[st] = diad_leg(P, cases, args.diad_levels, args.diad_steps, st)
            log(f"  {st}: {dl[st]['mlups']:.0f} MLUPS ({dl[st]['ms_per_coarse_step']:.2f} ms per coarse step), "
                f"cells per level {dl[st]['cells_per_level']}, build {dl[st]['build_ms']:.0f} ms")
        out["variants"]["diad_sphere"] = dl
    if args.closure_n:
        log(f"[bench] learned closure, {args.closure_n}^3 decaying turbulence")
        out["variants"]["closure"] = closure_leg(P, args.closure_n, log, args.closure_every,
                                                 args.closure_sphere_steps)
    if not args.no_cpu:
        log(f"[bench] cpu baseline (~{args.cpu_seconds:.0f} s)")
        out["cpu_baseline"] = cpu_baseline(cases, flags_keep, nu_keep, 0.17, shape, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

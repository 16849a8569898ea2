"""Benchmark: AMOEBA nonbonded energy+gradient evaluations per second on one MI355X.

Workload: water-20736 -- 6912 rigid AMOEBA-2003 waters (20736 atoms), seeded
random insertion (O-O >= 2.6 A), 59.1 A cubic box, PME grid 64^3 order 5,
Ewald cutoff 7 A (alpha 0.545), vdW cutoff 9 A, PCG to 1e-5 D.

A step is one full nonbonded energy+gradient evaluation (vdW, permanent PME,
PCG induced dipoles, reciprocal and real-space energies, torques ->
gradients) with coordinates already resident in HBM.  The mixed form is the
headline value; the FP64 form is timed beside it and the force relative RMS
error of mixed against FP64 is reported.  The CPU baseline is the FP64 C
oracle (same algorithm) on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "AMOEBA E+grad evals/s: mixed vs FP64 GPU at force rel RMS err <=1e-5"
N_WATERS, BOX, GRID, SEED = 6912, 59.1, (64, 64, 64), 2

# Algorithmic work of one launch of each kernel family (bytes that must cross
# HBM at least once, and FLOPs per pair counted from the reference formulas).
# s = bytes per real of the grid / pair arithmetic (4 mixed, 8 FP64).
FLOPS_PER_PAIR = {"k_real_energy": 1400.0, "k_perm_field": 220.0, "k_dip_field": 80.0, "k_vdw": 60.0}
PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}
PEAK_HBM_GBS = 8000.0


def kernel_family(name):
    for k in ("k_real_energy", "k_perm_field", "k_dip_field", "k_vdw", "k_ct_x", "k_ct_y", "k_ct_z",
              "k_spread_brick", "k_gather_ap", "k_gather", "k_filter", "k_plane"):
        if k in name:
            return k
    return name


def work_model(name, w):
    """(bytes, description, flops or None) of one launch.  w: n, K, s (real
    size), ewald_rows (entries of the Ewald rows of a launch = 2 x Ewald
    pairs), vdw_rows, nbricks, ext."""
    fam = kernel_family(name)
    n, K, s = w["n"], w["K"], w["s"]
    K3 = K ** 3
    cx = 2 * s  # complex grid point
    if fam == "k_dip_field":
        rec = 64 if s == 4 else 96
        rows = w["ewald_rows"]
        b = n * (rec + 4 + 48) + 4 * rows
        return b, f"{rec} B record + 4 B count + 48 B field per atom, 4 B per row entry ({rows} entries)", \
            FLOPS_PER_PAIR[fam] * rows
    if fam == "k_real_energy":
        rows = w["ewald_rows"]
        b = n * (32 + 96 + 2 * 48 + 48) + 4 * rows
        return b, "position, multipole, gradient/torque read-modify-write, dipoles per atom + row entries", \
            FLOPS_PER_PAIR[fam] * rows / 2
    if fam == "k_perm_field":
        rows = w["ewald_rows"]
        return n * (32 + 96 + 48) + 4 * rows, "position, multipole, two fields per atom + row entries", \
            FLOPS_PER_PAIR[fam] * rows
    if fam == "k_vdw":
        rows = w["vdw_rows"]
        return n * (32 + 4 + 24) + 4 * rows, "reduced site, class, gradient per atom + Verlet row entries", \
            FLOPS_PER_PAIR[fam] * rows
    if fam in ("k_ct_y", "k_ct_z") or (fam == "k_ct_x" and "0>" in name):
        b = 2 * K3 * cx + (K3 * s if fam == "k_ct_z" else 0)
        return b, "complex grid read + write" + (" + influence function" if fam == "k_ct_z" else ""), None
    if fam == "k_ct_x":
        nbuf = w["nbricks"] * w["ext"] ** 3 * (2 if "2>" in name else 1) * s
        return nbuf + K3 * cx, "brick buffer read + complex grid write", None
    if fam == "k_spread_brick":
        # the generic name covers 1 permanent and ~8 induced launches per
        # evaluation: the induced (two-part) launch is the model
        parts = 1 if "false" in name else 2
        per_atom = 60 * s + 12 + 4 + (48 if parts == 2 else 80)
        return n * per_atom + w["nbricks"] * w["ext"] ** 3 * parts * s, \
            "B-spline weights (grid type), grid start, atom index, dipoles / multipoles per atom + brick buffer write", \
            None
    if fam == "k_plane":
        if "true>" in name:
            return 2 * K3 * cx, "complex grid read + write (inverse y, x)", None
        nbuf = w["nbricks"] * w["ext"] ** 3 * (1 if ", 1, false" in name else 2) * s
        return nbuf + K3 * cx, "brick buffer read + complex grid write (combine, x, y)", None
    if fam == "k_gather_ap":
        return n * (60 * s + 12) + K3 * cx + n * 3 * 8 * 2 * 7, \
            "B-spline weights per atom + grid read + z, f, p, q (read / write) per atom", None
    if fam == "k_gather":
        return n * (60 * s + 12) + K3 * cx + n * 2 * 8 * (30 if "true" in name else 3), \
            "B-spline weights per atom + grid read + potentials written", None
    return None, None, None


def lanes_per_atom():
    """The library's lanes-per-atom macros (NB_LPA_PCG before NB_LPA, so the
    longer name is replaced first) and their values, from nb_kern.h."""
    vals = {}
    with open(os.path.join(ROOT, "forcefieldx_amd", "csrc", "nb_kern.h")) as f:
        for line in f:
            m = re.match(r"#define (NB_LPA\w*) (\d+)", line)
            if m:
                vals[m.group(1)] = m.group(2)
    return dict(sorted(vals.items(), key=lambda kv: -len(kv[0])))


def profile_key(name, prec="mixed"):
    """The kernel's name in the rocprofv3 summaries for a name the library's
    profiler reports (the launch expression, with template parameters by
    name): the mixed form's instantiation."""
    key = name.strip("()").split("(")[0]
    for macro, value in lanes_per_atom().items():
        key = key.replace(macro, value)
    real = "float" if prec == "mixed" else "double"
    key = key.replace("<T, IND>", f"<{real}, true>").replace("<T>", f"<{real}>").replace("<T, ", f"<{real}, ")
    return key


def roofline_for(name, avg_ms, w, prec):
    """HBM roofline of the kernel: algorithmic bytes of one launch / its average
    duration; pair kernels also carry their VALU rate against the FP32 / FP64
    vector peak."""
    secs = avg_ms * 1e-3
    b, model, flops = work_model(name, w)
    out = {"bound": "hbm", "achieved": None, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": None, "traffic": None,
           "kernel": name, "avg_launch_ms": avg_ms}
    if b is not None:
        gbs = b / secs / 1e9
        out.update({"achieved": gbs, "frac": gbs / PEAK_HBM_GBS, "algorithmic_bytes_per_launch": b,
                    "bytes_model": model})
    if flops is not None:
        real = "f32" if prec == "mixed" else "f64"
        tf = flops / secs / 1e12
        out.update({"valu_tflops": tf, "valu_peak_tflops": PEAK_TFLOPS[real], "valu_frac": tf / PEAK_TFLOPS[real],
                    "flops_per_launch": flops})
    return out


def measured_traffic(name):
    """HBM bytes per launch of the kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration),
    or None when the summary has no line for it."""
    path = os.path.join(ROOT, "profiles", "water-20736_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    key = profile_key(name)
    for line in open(path):
        if line.startswith("| `") and "|" in line[3:]:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if cells[0].strip("`") == key:
                try:
                    return float(cells[-1])
                except ValueError:
                    return None
    return None


def time_gpu(top, box, xyz, params, prec, steps, warmup, profile, graph=False, lam=None, valence=None):
    """Warm up, optionally find the dominant kernel with a 2-step fully timed
    calibration, then time `steps` evaluations.  Without graph mode, HIP events
    bracket only the launches of the dominant kernel inside the timed region
    (on the library's stream), so its average duration is measured live at
    negligible overhead.  In graph mode the timed region replays the recorded
    evaluation, and the dominant kernel is timed with HIP events over `steps`
    regular evaluations right after it (same kernels on the same data)."""
    from forcefieldx_amd.energy import ForceFieldEnergy
    ffe = ForceFieldEnergy(top, box, params, precision=prec, valence=valence)
    ffe.setCoordinates(xyz)
    h = ffe._handle
    if lam is not None:
        ffe.setLambda(lam)
    calib = {}
    dominant = None
    if profile:
        ffe._evaluate(True)
        h.set_profiling(True)
        for _ in range(2):
            ffe._evaluate(True)
        calib = h.profile()
        dominant = max(calib.items(), key=lambda kv: kv[1][0])[0]
        h.set_profiling(False)
    if graph:
        h.set_graph(True)
    elif profile:
        h.set_profiling(True, kernel_filter=dominant)
    for _ in range(warmup):
        ffe._evaluate(True)
    replays0 = h.graph_replays()
    t0 = time.perf_counter()
    for _ in range(steps):
        ffe._evaluate(True)
    t1 = time.perf_counter()
    replays = h.graph_replays() - replays0
    prof = {}
    if profile:
        if graph:
            h.set_graph(False)
            h.set_profiling(True, kernel_filter=dominant)
            for _ in range(steps):
                ffe._evaluate(True)
        prof = {"calibration": calib, "timed": h.profile(), "graph_replays": replays}
        h.set_profiling(False)
    last = dict(ffe._last)
    g = h.gradient()
    u, up = h.induced()
    extra = {"graph_replays": replays}
    if lam is not None:
        extra["dudl"] = ffe.getdEdL()
    ffe.destroy()
    return (t1 - t0) / steps, last, g, u, prof, extra


def time_gpu_perturbed(top, box, xyz, params, prec, steps, warmup, profile, graph=False, lam=None):
    """As time_gpu, but every evaluation sees new coordinates: a random walk
    with Gaussian steps of sigma 0.01 A per coordinate, whose frames are
    written to device memory before the timed region; each evaluation hands
    the library the next frame's device address (nb_set_positions_device, a
    device-to-device copy).  Both precisions see the same sequence, so the
    last gradients are comparable."""
    from forcefieldx_amd import ffi
    from forcefieldx_amd.energy import ForceFieldEnergy
    ffe = ForceFieldEnergy(top, box, params, precision=prec)
    h = ffe._handle
    if lam is not None:
        ffe.setLambda(lam)
    x0 = np.asarray(xyz, np.float64).reshape(-1)
    nprof = 2 if profile else 0
    nframes = nprof + warmup + steps
    frames = ffi.DeviceArray(x0.size, frames=nframes)
    # frames [0, nprof): the profiling evaluations' own walk; then warm-up and
    # timed frames from the same seed in both forms
    for seed, first, count in ((4321, 0, nprof), (1234, nprof, warmup + steps)):
        rng = np.random.default_rng(seed)
        x = x0.copy()
        for k in range(first, first + count):
            x += rng.normal(0.0, 0.01, x.shape)
            frames.upload(x, k)
    cur = [0]

    def step():
        h.set_positions_device(frames.ptr(cur[0]))
        cur[0] += 1
        ffe._evaluate(True)

    ffe._x = x0.copy()
    h.set_positions(ffe._x)
    calib, dominant = {}, None
    if profile:
        ffe._evaluate(True)
        h.set_profiling(True)
        for _ in range(2):
            step()
        calib = h.profile()
        dominant = max(calib.items(), key=lambda kv: kv[1][0])[0]
        h.set_profiling(False)
    if graph:
        h.set_graph(True)
    elif profile:
        h.set_profiling(True, kernel_filter=dominant)
    for _ in range(warmup):
        step()
    replays0 = h.graph_replays()
    ffi.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    t1 = time.perf_counter()
    replays = h.graph_replays() - replays0
    last = dict(ffe._last)
    g = h.gradient()
    u, up = h.induced()
    prof = {}
    if profile:
        if graph:
            h.set_graph(False)
            h.set_profiling(True, kernel_filter=dominant)
            cur[0] = nframes - steps
            for _ in range(steps):
                step()
        prof = {"calibration": calib, "timed": h.profile(), "graph_replays": replays}
        h.set_profiling(False)
    extra = {"graph_replays": replays}
    if lam is not None:
        extra["dudl"] = ffe.getdEdL()
    ffe.destroy()
    frames.close()
    return (t1 - t0) / steps, last, g, u, prof, extra


def time_batch(systems, params, prec, steps, warmup, kernel=True):
    """One step = nb_evaluate_batch over every cell.  kernel: the single
    persistent launch of the batch kernel (the default path); otherwise the
    separate path (NB_BATCH_SEPARATE), whose recorded graphs of all cells are
    launched together, each on its own stream.  Returns the mean step time,
    the energies, gradients, graph replays, PCG iterations and the batch
    kernel's own launch durations (HIP events around each timed launch)."""
    from forcefieldx_amd import ffi
    hs = []
    for xyz, box, top, grid in systems:
        h = ffi.Handle(ffi.params_from(params, grid, prec), top)
        h.set_cell(box)
        h.set_positions(xyz)
        if not kernel:
            h.set_graph(True)
        hs.append(h)
    flags = ffi.NB_GRAD | (0 if kernel else ffi.NB_BATCH_SEPARATE)
    kms = []
    try:
        for _ in range(warmup):
            ffi.evaluate_batch(hs, flags)
        r0 = sum(h.graph_replays() for h in hs)
        t0 = time.perf_counter()
        for _ in range(steps):
            es, rc = ffi.evaluate_batch(hs, flags)
            if rc != ffi.NB_OK:
                raise RuntimeError(f"batch evaluation failed: {rc}")
            if kernel:
                kms.append(ffi.batch_config()["kernel_ms"])
        t1 = time.perf_counter()
        replays = sum(h.graph_replays() for h in hs) - r0
        grads = [h.gradient().reshape(-1) for h in hs]
        iters = [e["pcg_iters"] for e in es]
        cfg = ffi.batch_config() if kernel else None
    finally:
        for h in hs:
            h.close()
    return (t1 - t0) / steps, es, np.concatenate(grads), replays, iters, kms, cfg


def batch_bytes(systems, iters, s):
    """Algorithmic HBM bytes of one sweep, summed over the cells: per cell
    12 s K^3 (I + 3) for the grid passes, N (22 s + 48)(2 I + 5) for the atom
    records and fields of the real-space passes, N (4 s + 24) for vdW,
    N s (73 + 12 I) for the spline and potential streams and 384 N I for the
    FP64 PCG vectors, with I the cell's PCG iterations."""
    tot = 0.0
    for (xyz, box, top, grid), I in zip(systems, iters):
        n, K3 = top.n, grid[0] * grid[1] * grid[2]
        tot += (12 * s * K3 * (I + 3) + n * (22 * s + 48) * (2 * I + 5) + n * (4 * s + 24) + n * s * (73 + 12 * I)
                + 384 * n * I)
    return tot


def batch_main(args, systems, params):
    from forcefieldx_amd import ffi
    t64, e64, g64, _, it64, k64, cfg64 = time_batch(systems, params, ffi.NB_PREC_FP64, args.steps, args.warmup)
    tmx, emx, gmx, _, itmx, kmx, cfgmx = time_batch(systems, params, ffi.NB_PREC_MIXED, args.steps, args.warmup)
    tsep = gsep = rsep = None
    if not args.no_separate:
        sep_steps = min(args.steps, 3)
        tsep, _, gsep, rsep, _, _, _ = time_batch(systems, params, ffi.NB_PREC_MIXED, sep_steps, 1, kernel=False)
    natoms = sum(s[2].n for s in systems)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import dataclasses
        import pyoracle
        pyoracle.build()
        nth = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        t0 = time.perf_counter()
        done = 0
        for xyz, box, top, grid in systems:
            pyoracle.evaluate(pyoracle.System(xyz, box, top), dataclasses.replace(params, grid=grid), nthreads=nth)
            done += 1
            if time.perf_counter() - t0 > 12.0:
                break
        dt = time.perf_counter() - t0
        cpu = {"value": done / dt / len(systems), "unit": "sweeps/s", "cores": nth, "kind": "port",
               "sample": f"{done} of the {len(systems)} cells evaluated one after the other by the FP64 oracle "
                         f"({dt:.1f} s), scaled to a full sweep"}
    kavg = float(np.mean(kmx))
    bytes_sweep = batch_bytes(systems, itmx, 4)
    gbs = bytes_sweep / (kavg * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
            "traffic": None, "kernel": "k_batch", "avg_launch_ms": kavg, "algorithmic_bytes_per_launch": bytes_sweep,
            "bytes_model": "sum over cells of the whole-evaluation byte count (grid passes, atom records and "
                           "fields, vdW, spline and potential streams, FP64 PCG vectors) at 4-byte reals",
            "timed_with": "HIP events around each batch-kernel launch of the timed region (library stream)"}
    out = {
        "metric": METRIC, "value": 1.0 / tmx, "unit": "sweeps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": tmx * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "mixed: f32 pair arithmetic, f64 accumulation/PME/PCG", "data": "synthetic",
        "config": {"workload": f"crystals-512: {WORKLOADS['crystals-512']}", "cells": len(systems), "atoms": natoms,
                   "launch": "one persistent launch of the batch kernel per sweep",
                   "kernel_shape": {k: cfgmx[k] for k in ("threads", "lanes", "workgroups")},
                   "parallelism": "single GPU"},
        "cell_evals_per_s": len(systems) / tmx,
        "fp64_ms_per_step": t64 * 1e3, "mixed_over_fp64": t64 / tmx,
        "kernel_ms": {"mixed": kavg, "fp64": float(np.mean(k64))},
        "separate_path_ms_per_step": tsep * 1e3 if tsep else None,
        "separate_path_graph_replays": rsep,
        "kernel_over_separate_path": tsep / tmx if tsep else None,
        "force_rel_rms_err": float(np.linalg.norm(gmx - g64) / np.linalg.norm(g64)),
        "separate_vs_kernel_force_rel_err": (float(np.linalg.norm(gsep - gmx) / np.linalg.norm(gmx))
                                             if gsep is not None else None),
        "pcg_iters_mean": {"mixed": float(np.mean(itmx)), "fp64": float(np.mean(it64))},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = out["value"] / cpu["value"]
    print(json.dumps(out))


def accuracy_probe(top, box, xyz, params, grid, lam=None, valence=None):
    """Errors of the mixed form against the FP64 form on one evaluation of the
    given coordinates with the strain derivative (outside any timed region):
    max and RMS relative errors of forces, induced dipoles and virial, and the
    energy difference."""
    from forcefieldx_amd import ffi
    res = {}
    for prec in (ffi.NB_PREC_FP64, ffi.NB_PREC_MIXED):
        h = ffi.Handle(ffi.params_from(params, grid, prec), top)
        try:
            h.set_cell(box)
            if valence is not None:
                h.set_valence(valence)
            if lam is not None:
                h.set_lambda(lam)
            h.set_positions(np.asarray(xyz, np.float64).reshape(-1))
            e, _ = h.evaluate(ffi.NB_GRAD | ffi.NB_VIRIAL)
            res[prec] = (e, h.gradient().reshape(-1), h.induced()[0].reshape(-1), h.virial().reshape(-1))
        finally:
            h.close()
    (e64, g64, u64, w64), (emx, gmx, umx, wmx) = res[ffi.NB_PREC_FP64], res[ffi.NB_PREC_MIXED]

    def errs(a, b):
        return {"rms": float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)),
                "max": float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))}
    return {"force": errs(gmx, g64), "induced": errs(umx, u64), "virial": errs(wmx, w64),
            "energy_abs": float(emx["total"] - e64["total"]),
            "energy_rel": float(abs(emx["total"] - e64["total"]) / max(abs(e64["total"]), 1e-300)),
            "note": "one evaluation with NB_GRAD | NB_VIRIAL of the workload's starting coordinates"}


def cpu_baseline(top, box, xyz, params, workload, lam=None, budget_s=12.0, valence=None):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.build()
    nth = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    import dataclasses
    if lam is not None:
        params = dataclasses.replace(params, lambda_value=lam)
    s = pyoracle.System(xyz, box, top)
    t0 = time.perf_counter()
    count = 0
    while True:
        e, g, lg = pyoracle.evaluate_lambda(s, params, nthreads=nth)
        if valence is not None:
            ev, gv = pyoracle.valence(valence, xyz)
            e = dict(e, total=e["total"] + ev["total"])
            g = g + gv
        count += 1
        if time.perf_counter() - t0 > budget_s or count >= 50:
            break
    dt = time.perf_counter() - t0
    return {"value": count / dt, "unit": "evals/s", "cores": nth, "kind": "port",
            "sample": f"{count} full FP64 E+G evaluations of the same {workload} workload "
                      f"(oracle/amoeba_oracle.c, OpenMP" + (", valence terms by oracle/valence_oracle.c"
                                                            if valence is not None else "") +
                      f", {dt:.1f} s)"}, e, g


WORKLOADS = {
    # name: (description, builder) -- the first is the headline configuration
    "water-20736": "6912 AMOEBA-2003 waters, 59.1 A cubic box, PME 64^3 order 5, Ewald 7 A, vdW 9 A, PCG 1e-5 D",
    "water-288000": "96000 AMOEBA-2003 waters, 144.3 A cubic box, PME 144^3 order 5, Ewald 7 A, vdW 9 A, PCG 1e-5 D, "
                    "each step on the next frame of a random walk of sigma 0.01 A steps, the frames resident in "
                    "device memory before the timed region",
    "chain-23557": "2500-atom random-walk chain with random multipoles in 7019 waters, 62.2 A box, PME 64^3, "
                   "20 soft atoms at lambda 0.5: OST electrostatics + vdW softcore with dU/dL",
    "water-648": "216 AMOEBA-2003 waters, 18.643 A cubic box, PME 24^3, PCG 1e-6 D",
    "dhfr-23558": "the reference's DHFR benchmark (EnergyTest): 23,558 atoms of AMOEBA-BIO-2009 protein and water, "
                  "62.23 A cubic box, PME 80^3 order 5, Ewald 6.7 A (alpha 0.57), vdW 9 A, SOR to 1e-6 D as its "
                  "key file asks, plus the 47,913 valence terms; coordinates and parameters from the Tinker "
                  "fixture in tests/golden/dhfr",
    "ubiquitin-9737": "the reference's ubiquitin benchmark (EnergyTest): 9,737 atoms of AMOEBA-BIO-2009 protein and "
                      "water, 54.99 x 41.91 x 41.91 A cell, PME 70 x 54 x 54 order 5, Ewald 6.7 A (alpha 0.57), "
                      "vdW 9 A, PCG to 1e-6 D, plus the 20,920 valence terms; coordinates and parameters from "
                      "the Tinker fixture in tests/golden/ubiquitin",
    "ubiquitin-md": "velocity-Verlet dynamics of the reference's ubiquitin benchmark structure (9,737 atoms, "
                    "AMOEBA-BIO-2009 valence and nonbonded terms, PME 70 x 54 x 54, PCG to 1e-6 D) at 1 fs steps, "
                    "ASPC SCF predictor, each step (drift, evaluation, kick) replayed as one HIP graph; a step is one "
                    "MD step",
    "crystals-512": "512 independent P1 triclinic cells (lengths 8-20 A, angles 70-110 deg) of small molecules "
                    "with random multipoles, per-cell PME grids 16-32, Ewald and vdW 6 A, PCG 1e-5 D, image sums "
                    "in cells thinner than twice the cutoff; a step is one sweep over all cells (nb_evaluate_batch)",
}


def build_workload(name):
    from forcefieldx_amd.forcefield import NonbondedParams
    from forcefieldx_amd.systems import chain_in_water, crystal_batch, water_box
    lam = None
    if name == "water-20736":
        xyz, box, top = water_box(N_WATERS, BOX, seed=SEED)
        params, grid = NonbondedParams(vdw_cutoff=9.0, grid=GRID, polar_eps=1e-5), GRID
    elif name == "water-288000":
        xyz, box, top = water_box(96000, 144.3, seed=SEED)
        grid = (144, 144, 144)
        params = NonbondedParams(vdw_cutoff=9.0, grid=grid, polar_eps=1e-5)
    elif name == "chain-23557":
        xyz, box, top = chain_in_water(2500, 7019, 62.2, seed=7, n_soft=20)
        grid = (64, 64, 64)
        params = NonbondedParams(vdw_cutoff=9.0, grid=grid, polar_eps=1e-5, elec_lambda=True)
        lam = 0.5
    elif name == "crystals-512":
        systems = crystal_batch(512, seed=SEED)
        params = NonbondedParams(ewald_cutoff=6.0, vdw_cutoff=6.0, polar_eps=1e-5)
        return systems, None, None, params, None, None, None
    elif name in ("dhfr-23558", "ubiquitin-9737"):
        from forcefieldx_amd import tinker
        stem = name.split("-")[0]
        gold = os.path.join(ROOT, "tests", "golden", stem)
        xyz, box, top, params, valence = tinker.load(
            os.path.join(gold, f"{stem}.xyz.gz"), os.path.join(gold, f"{stem}.key"),
            os.path.join(gold, f"amoeba_bio_2009_{stem}.prm.gz"), valence=True)
        return xyz, box, top, params, params.grid, lam, valence
    elif name == "water-648":
        xyz, box, top = water_box(216, 18.643, seed=1)
        grid = (24, 24, 24)
        params = NonbondedParams(vdw_cutoff=9.0, grid=grid, polar_eps=1e-6)
    else:
        raise SystemExit(f"unknown workload {name}; one of {', '.join(WORKLOADS)}")
    return xyz, box, top, params, grid, lam, None


def md_main(args):
    """Steps per second of device-resident velocity-Verlet dynamics."""
    from forcefieldx_amd import ffi, tinker
    from forcefieldx_amd.dynamics import MolecularDynamics
    from forcefieldx_amd.energy import ForceFieldEnergy
    gold = os.path.join(ROOT, "tests", "golden", "ubiquitin")
    xyz, box, top, params, valence = tinker.load(os.path.join(gold, "ubiquitin.xyz.gz"),
                                                 os.path.join(gold, "ubiquitin.key"),
                                                 os.path.join(gold, "amoeba_bio_2009_ubiquitin.prm.gz"), valence=True)
    ff = tinker.read_prm(os.path.join(gold, "amoeba_bio_2009_ubiquitin.prm.gz"))
    mass = tinker.masses(tinker.read_xyz(os.path.join(gold, "ubiquitin.xyz.gz")).types, ff)
    dt_fs = 1.0
    res = {}
    for name, prec in (("fp64", ffi.NB_PREC_FP64), ("mixed", ffi.NB_PREC_MIXED)):
        ffe = ForceFieldEnergy(top, box, params, precision=prec, valence=valence)
        md = MolecularDynamics(ffe, mass)
        md.init(xyz, md.initVelocities(298.15, seed=SEED))
        r0 = md.dynamic(0, dt_fs)
        md.dynamic(max(args.warmup, 8), dt_fs)     # fills the predictor's history
        t0 = time.perf_counter()
        r = md.dynamic(args.steps, dt_fs)
        t = (time.perf_counter() - t0) / args.steps
        res[name] = dict(s_per_step=t, pcg_iters_per_step=r["pcg_iters"] / max(r["steps"], 1),
                         total_drift=r["total"] - r0["total"], kinetic=r["kinetic"],
                         replays=ffe._handle.graph_replays())
        ffe.destroy()
    tm = res["mixed"]["s_per_step"]
    cpu = None
    if not args.no_cpu:
        cpu, _, _ = cpu_baseline(top, box, xyz, params, "ubiquitin-md", valence=valence)
        cpu = dict(cpu, unit="steps/s", sample=cpu["sample"] + "; one energy and gradient per MD step, the "
                   "integrator's own work negligible beside it")
    out = {
        "metric": "AMOEBA MD steps/s (velocity Verlet, 1 fs)",
        "value": 1.0 / tm,
        "unit": "steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": tm * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "mixed: f32 pair arithmetic, f64 accumulation/PME/PCG/integration",
        "data": "the reference's own test structure (tests/golden/ubiquitin), Maxwell-Boltzmann velocities at 298 K",
        "config": {"workload": f"ubiquitin-md: {WORKLOADS['ubiquitin-md']}", "atoms": top.n, "grid": list(params.grid),
                   "launch": "one HIP graph replay per MD step", "parallelism": "single GPU"},
        "ns_per_day": 86400.0 / tm * dt_fs * 1e-6,
        "fp64_steps_per_s": 1.0 / res["fp64"]["s_per_step"],
        "fp64_ns_per_day": 86400.0 / res["fp64"]["s_per_step"] * dt_fs * 1e-6,
        "pcg_iters_per_step": {k: v["pcg_iters_per_step"] for k, v in res.items()},
        "total_energy_drift": {k: v["total_drift"] for k, v in res.items()},
        "graph_replays": {k: v["replays"] for k, v in res.items()},
        "roofline": None,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = out["value"] / cpu["value"]
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="water-20736", choices=list(WORKLOADS))
    ap.add_argument("--no-graph", action="store_true", help="time the regular launch path instead of graph replays")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-separate", action="store_true",
                    help="crystals-512: skip timing the separate (graph per cell) path beside the batch kernel")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 256 MiB device-to-device copy of known size (reference point for "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters); not timed")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(see the sharding note in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from forcefieldx_amd import build as nbbuild
    from forcefieldx_amd import ffi
    if not os.path.exists(ffi.LIB_PATH):
        nbbuild.build()
    if args.copy_calibration:
        import torch
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.workload == "ubiquitin-md":
        return md_main(args)
    xyz, box, top, params, grid, lam, valence = build_workload(args.workload)
    if args.workload == "crystals-512":
        return batch_main(args, xyz, params)
    graph = not args.no_graph
    perturb = args.workload == "water-288000"
    runner = time_gpu_perturbed if perturb else time_gpu

    vk = {"valence": valence} if valence is not None else {}
    t64, e64, g64, u64, prof64, x64 = runner(top, box, xyz, params, ffi.NB_PREC_FP64, args.steps, args.warmup, False,
                                             graph, lam, **vk)
    tmx, emx, gmx, umx, profmx, xmx = runner(top, box, xyz, params, ffi.NB_PREC_MIXED, args.steps, args.warmup,
                                             not args.no_profile, graph, lam, **vk)
    frc_rms = float(np.linalg.norm(gmx - g64) / np.linalg.norm(g64))
    frc_max = float(np.max(np.abs(gmx - g64)) / np.max(np.abs(g64)))
    ind_rms = float(np.linalg.norm(umx - u64) / max(np.linalg.norm(u64), 1e-300))
    accuracy = accuracy_probe(top, box, xyz, params, grid, lam, valence)

    roof = None
    if profmx:
        name, (tot, cnt) = max(profmx["timed"].items(), key=lambda kv: kv[1][0])
        nbk = max(2, (grid[0] + 7) // 8)
        w = {"n": top.n, "K": grid[0], "s": 4, "ewald_rows": 2 * (int(emx["n_perm"]) - top.n),
             "vdw_rows": 2 * int(emx["n_vdw"]), "nbricks": nbk ** 3, "ext": -(-grid[0] // nbk) + 4}
        roof = roofline_for(name, tot / cnt, w, "mixed")
        kernel_steps = args.steps
        roof["launches_per_step"] = cnt / kernel_steps
        roof["share_of_step"] = tot / (tmx * 1e3 * kernel_steps)
        roof["traffic"] = measured_traffic(name) if args.workload == "water-20736" else None
        roof["timed_with"] = ("HIP events around each launch of this kernel in the timed region" if not graph else
                              "HIP events around each launch of this kernel over the same number of regular "
                              "evaluations right after the graph-replay timed region")
        # the other kernels of the PCG iteration's critical path (spread, FFT
        # planes, gather) and the real-space field, from the calibration pass
        fams = []
        for kname, (ktot, kcnt) in sorted(profmx["calibration"].items(), key=lambda kv: -kv[1][0]):
            if kernel_family(kname) in ("k_spread_brick", "k_plane", "k_ct_x", "k_ct_y", "k_ct_z", "k_gather_ap",
                                        "k_dip_field") and kcnt > 0:
                r = roofline_for(kname, ktot / kcnt, w, "mixed")
                r["traffic"] = measured_traffic(kname) if args.workload == "water-20736" else None
                r["launches_per_step"] = kcnt / 2
                fams.append({k: r.get(k) for k in ("kernel", "avg_launch_ms", "achieved", "frac", "traffic",
                                                   "launches_per_step", "valu_frac")})
        roof["path_kernels"] = fams[:8]

    cpu = None
    e_cpu = None
    if not args.no_cpu:
        cpu, e_cpu, g_cpu = cpu_baseline(top, box, xyz, params, args.workload, lam, valence=valence)

    out = {
        "metric": METRIC,
        "value": 1.0 / tmx,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tmx * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "mixed: f32 pair arithmetic, f64 accumulation/PME/PCG",
        "data": (f"the reference's own test structure (tests/golden/{args.workload.split('-')[0]})"
                 if args.workload in ("dhfr-23558", "ubiquitin-9737") else "synthetic"),
        "config": {"workload": f"{args.workload}: {WORKLOADS[args.workload]}", "atoms": top.n, "grid": list(grid),
                   "launch": ("regular launches (SOR solves are not recorded)" if params.scf_algorithm == 1 else
                              "HIP graph replay" if graph else "regular launches"),
                   "parallelism": "single GPU"},
        "fp64_evals_per_s": 1.0 / t64,
        "fp64_ms_per_step": t64 * 1e3,
        "mixed_over_fp64": t64 / tmx,
        "force_rel_rms_err": frc_rms,
        "force_rel_max_err": frc_max,
        "induced_rel_rms_err": ind_rms,
        "energy_mixed_minus_fp64": emx["total"] - e64["total"],
        "mixed_vs_fp64": accuracy,
        "energy_fp64": e64["total"],
        "pcg_iters": {"mixed": emx["pcg_iters"], "fp64": e64["pcg_iters"]},
        "pairs": {"ewald": int(e64["n_perm"]) - top.n, "vdw": int(e64["n_vdw"])},
        "graph_replays": {"mixed": xmx["graph_replays"], "fp64": x64["graph_replays"]},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if lam is not None:
        out["dudl"] = {"mixed": xmx["dudl"], "fp64": x64["dudl"]}
    if cpu:
        out["gpu_over_cpu"] = out["value"] / cpu["value"]
        if not perturb:
            out["energy_fp64_minus_oracle"] = e64["total"] - e_cpu["total"]
    if profmx:
        out["kernel_ms_per_step_calibration"] = {k: round(v[0] / 2, 4) for k, v in
                                                 sorted(profmx["calibration"].items(), key=lambda kv: -kv[1][0])}
    print(json.dumps(out))


if __name__ == "__main__":
    main()

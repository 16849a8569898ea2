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
              "k_spread_brick", "k_gather", "k_filter"):
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
        parts = 2 if "true" in name else 1
        per_atom = 480 + 12 + 4 + (48 if parts == 2 else 80)
        return n * per_atom + w["nbricks"] * w["ext"] ** 3 * parts * s, \
            "B-spline weights, grid start, dipoles / multipoles per atom + brick buffer write", None
    if fam == "k_gather":
        return n * (480 + 12) + K3 * cx + n * 2 * 8 * (30 if "true" in name else 3), \
            "B-spline weights per atom + grid read + potentials written", None
    return None, None, None


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
    key = name.strip("()").split("(")[0]
    for line in open(path):
        if line.startswith("| `") and "|" in line[3:]:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if cells[0].strip("`") == key:
                try:
                    return float(cells[-1])
                except ValueError:
                    return None
    return None


def time_gpu(top, box, xyz, params, prec, steps, warmup, profile):
    """Warm up, optionally find the dominant kernel with a 2-step fully timed
    calibration, then time `steps` evaluations.  In the timed region HIP events
    bracket only the launches of the dominant kernel (on the library's
    stream), so the kernel's average duration is measured live at negligible
    overhead."""
    from forcefieldx_amd.energy import ForceFieldEnergy
    ffe = ForceFieldEnergy(top, box, params, precision=prec)
    ffe.setCoordinates(xyz)
    h = ffe._handle
    for _ in range(warmup):
        ffe._evaluate(True)
    calib = {}
    if profile:
        h.set_profiling(True)
        for _ in range(2):
            ffe._evaluate(True)
        calib = h.profile()
        dominant = max(calib.items(), key=lambda kv: kv[1][0])[0]
        h.set_profiling(True, kernel_filter=dominant)
    t0 = time.perf_counter()
    for _ in range(steps):
        ffe._evaluate(True)
    t1 = time.perf_counter()
    prof = {}
    if profile:
        prof = {"calibration": calib, "timed": h.profile()}
        h.set_profiling(False)
    last = dict(ffe._last)
    g = h.gradient()
    u, up = h.induced()
    ffe.destroy()
    return (t1 - t0) / steps, last, g, u, prof


def cpu_baseline(top, box, xyz, params, budget_s=12.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.build()
    nth = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    s = pyoracle.System(xyz, box, top)
    t0 = time.perf_counter()
    count = 0
    while True:
        e, g, u, up = pyoracle.evaluate(s, params, nthreads=nth)
        count += 1
        if time.perf_counter() - t0 > budget_s or count >= 50:
            break
    dt = time.perf_counter() - t0
    return {"value": count / dt, "unit": "evals/s", "cores": nth, "kind": "port",
            "sample": f"{count} full FP64 E+G evaluations of the same water-20736 workload "
                      f"(oracle/amoeba_oracle.c, OpenMP, {dt:.1f} s)"}, e, g


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
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
    from forcefieldx_amd.forcefield import NonbondedParams
    from forcefieldx_amd.systems import water_box
    if not os.path.exists(ffi.LIB_PATH):
        nbbuild.build()
    if args.copy_calibration:
        import torch
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    xyz, box, top = water_box(N_WATERS, BOX, seed=SEED)
    params = NonbondedParams(vdw_cutoff=9.0, grid=GRID, polar_eps=1e-5)

    t64, e64, g64, u64, prof64 = time_gpu(top, box, xyz, params, ffi.NB_PREC_FP64, args.steps, args.warmup, False)
    tmx, emx, gmx, umx, profmx = time_gpu(top, box, xyz, params, ffi.NB_PREC_MIXED, args.steps, args.warmup,
                                          not args.no_profile)
    frc_rms = float(np.linalg.norm(gmx - g64) / np.linalg.norm(g64))
    frc_max = float(np.max(np.abs(gmx - g64)) / np.max(np.abs(g64)))
    ind_rms = float(np.linalg.norm(umx - u64) / np.linalg.norm(u64))

    roof = None
    if profmx:
        name, (tot, cnt) = max(profmx["timed"].items(), key=lambda kv: kv[1][0])
        nbk = max(2, (GRID[0] + 7) // 8)
        w = {"n": top.n, "K": GRID[0], "s": 4, "ewald_rows": 2 * (int(emx["n_perm"]) - top.n),
             "vdw_rows": 2 * int(emx["n_vdw"]), "nbricks": nbk ** 3, "ext": -(-GRID[0] // nbk) + 4}
        roof = roofline_for(name, tot / cnt, w, "mixed")
        roof["share_of_step"] = tot / (tmx * 1e3 * args.steps)
        roof["launches_per_step"] = cnt / args.steps
        roof["traffic"] = measured_traffic(name)

    cpu = None
    e_cpu = None
    if not args.no_cpu:
        cpu, e_cpu, g_cpu = cpu_baseline(top, box, xyz, params)

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
        "data": "synthetic",
        "config": {"workload": "water-20736: 6912 AMOEBA-2003 waters, 59.1 A cubic box, PME 64^3 order 5, "
                               "Ewald 7 A, vdW 9 A, PCG 1e-5 D", "atoms": top.n, "grid": list(GRID),
                   "seed": SEED, "parallelism": "single GPU"},
        "fp64_evals_per_s": 1.0 / t64,
        "fp64_ms_per_step": t64 * 1e3,
        "mixed_over_fp64": t64 / tmx,
        "force_rel_rms_err": frc_rms,
        "force_rel_max_err": frc_max,
        "induced_rel_rms_err": ind_rms,
        "energy_mixed_minus_fp64": emx["total"] - e64["total"],
        "energy_fp64": e64["total"],
        "pcg_iters": {"mixed": emx["pcg_iters"], "fp64": e64["pcg_iters"]},
        "pairs": {"ewald": int(e64["n_perm"]) - top.n, "vdw": int(e64["n_vdw"])},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = out["value"] / cpu["value"]
        out["energy_fp64_minus_oracle"] = e64["total"] - e_cpu["total"]
    if profmx:
        out["kernel_ms_per_step_calibration"] = {k: round(v[0] / 2, 4) for k, v in
                                                 sorted(profmx["calibration"].items(), key=lambda kv: -kv[1][0])}
    print(json.dumps(out))


if __name__ == "__main__":
    main()

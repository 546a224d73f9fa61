#!/usr/bin/env python3
"""Throughput of the ScLETD step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload beta_alpha_192]

A step is one pass of the reference's time-loop body (check_soln_new, then the
ETD step with transfer, elasticity, nonlinear term, mode products and phi
update) over all fields of the workload, inputs resident in HBM.  Prints one
JSON line: value = simulated time per wall second (dt * K / seconds), the
dominant kernel's roofline (FP64 MFMA mode products, timed with HIP events on
the library's stream) and the CPU baseline (the FP64 oracle, on this host's
cores, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "energy-stability_amd"))

METRIC = "sim-time/s at 192^3x12 vars: mixed-prec+learned dt vs plain FP64 at fixed err"
# HBM bytes per mode-product launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# (scaled against a copy kernel of known size) for the default workload
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_modeprod_dl_192cube_13fields_hbm_counters.md")


def measured_traffic(workload):
    if workload != "beta_alpha_192" or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if "traffic_bytes_per_launch_avg:" in line:
            return float(line.split("traffic_bytes_per_launch_avg:")[1].strip(" *\n"))
    return None
PEAK_FP64_MFMA = 78.6e12  # MI355X dense FP64 matrix, FLOP/s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="beta_alpha_192")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one plain device copy of this size before the timed steps (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs one subdomain on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pyetd
    from pyetd import workloads as W

    prm = getattr(W, args.workload)()
    n, V = prm["n"], prm["nac"] + prm["nch"]
    N = n - 2 * prm["g"]
    g = pyetd.Etd.from_dict(prm)
    for v, f in enumerate(W.initial_fields(prm)):
        g.set_field(v, f)
    if args.warmup:
        g.time_steps(args.warmup, with_diag=True)
    if args.calib_copy_mib:
        g.calib_copy(args.calib_copy_mib << 20)  # plain copy kernel, exactly this many bytes each way
    ms, launches = g.time_steps(args.steps, with_diag=True)
    total_s = ms[0] / 1e3
    K = args.steps
    dt = prm["dt"]
    value = dt * K / total_s
    stages = 2 if prm["etd2"] else 1
    # mode products per step: (2V forward + V backward) x 3 for ETD1; ETD2 adds 6V
    nprod = 9 * V + (6 * V if prm["etd2"] else 0)
    flops_step = nprod * 2.0 * n ** 4
    mp_s = ms[1] / 1e3
    achieved = flops_step * K / mp_s if mp_s > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "sim-time/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms[0] / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded noise initial fields, seeded cubic-rotation eigenstrains)",
        "config": {
            "workload": f"{prm['name']}: plain FP64 ETD{stages} drop-in, {N}^3 interior (n={n}), "
                        f"{prm['nac']} AC + {prm['nch']} CH fields, elastic, dt={dt}",
            "grid": N, "extended_edge": n, "fields": V, "dt": dt, "precision": "fp64", "dt_control": "fixed",
        },
        "grid_point_variant_steps_per_s": N ** 3 * V * K / total_s,
        "breakdown_ms_per_step": {
            "mode_products": ms[1] / K, "elastic_fft": ms[2] / K, "nonlinear": ms[3] / K, "phi_scaling": ms[4] / K,
            "energy": ms[5] / K, "halo": ms[6] / K,
        },
        "roofline": {
            "kernel": "k_modeprod_dl<13,4,2,2> (FP64 MFMA 4x4x4 mode products, direct-to-LDS pipeline)",
            "bound": "mfma",
            "achieved": achieved / 1e12,
            "peak": PEAK_FP64_MFMA / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_MFMA,
            "traffic": measured_traffic(args.workload),
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1, see profiles/)",
            "launches": launches[0],
            "flops_per_step": flops_step,
        },
    }
    g.close()
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        o = O.Oracle(prm)
        for v, f in enumerate(W.initial_fields(prm)):
            o.set_field(v, f)
        t0 = time.time()
        for _ in range(args.cpu_steps):
            o.diag()
            o.step()
        cs = time.time() - t0
        out["cpu_baseline"] = {
            "value": dt * args.cpu_steps / cs,
            "unit": "sim-time/s",
            "cores": O.threads(),
            "kind": "port",
            "sample": f"{args.cpu_steps} full step(s) of the same workload (check_soln_new + ETD{stages} step, "
                      f"{N}^3 x {V} fields) with the FP64 C oracle, OpenMP, {cs:.1f} s",
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

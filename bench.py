#!/usr/bin/env python3
"""Simulated time per wall second of the ScLETD step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload beta_alpha_192]

Workload: 192^3 interior (n = 196) x (12 AC + 1 CH) beta -> alpha fields,
cubic elasticity, ETD1, dt = 0.005, the 1000-step horizon t = 5.  Inputs are
resident in HBM before every timed region.

- value: the headline path, mixed precision (FP32 fields, MFMA mode products
  as FP16 split-2 by default or BF16 split-3 with --precision 1, FP32
  accumulation, FP64 energy) with adaptive steps proposed by the learned
  controller (step doubling, tolerance 1e-4, energy-decrease acceptance),
  integrating the whole horizon t = 0 -> 5: value = 5 / wall seconds.
- plain_fp64: the drop-in path (FP64, fixed dt = 0.005): K timed steps after
  W warmup steps, diagnostics included; sim-time/s = dt K / seconds.
- accuracy: the fixed-error check, max|d eta| of the headline run against the
  plain FP64 path after its 1000 steps (the plain FP64 GPU path equals the FP64
  oracle within 1e-10: tests/test_gpu_parity.py), the relative energy error,
  monotone energy; and the mixed-precision fixed-dt report every 100 steps.
- roofline: the dominant kernel of the headline path (mixed precision, timed
  at fixed dt with HIP events on the library's stream); traffic from the
  committed rocprofv3 counters.
- cpu_baseline: the FP64 C oracle on this host, one full step.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "energy-stability_amd"))

METRIC = "sim-time/s at 192^3x12 vars: mixed-prec+learned dt vs plain FP64 at fixed err"
PEAK_FP64_MFMA = 78.6e12  # MI355X dense FP64 matrix, FLOP/s
PEAK_BF16_MFMA = 2.5e15   # dense BF16 / FP16 matrix
PEAK_HBM = 8.0e12         # HBM3E, B/s
CONTROLLER = os.path.join(ROOT, "energy-stability_amd", "controller", "mlp.npz")
PROFILES = os.path.join(ROOT, "profiles")


def measured_traffic(name):
    """HBM bytes per launch from a committed counter summary, or None."""
    path = os.path.join(PROFILES, name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if "traffic_bytes_per_launch_avg:" in line:
            return float(line.split("traffic_bytes_per_launch_avg:")[1].strip(" *\n"))
    return None


def fresh(pyetd, W, prm):
    g = pyetd.Etd.from_dict(prm)
    for v, f in enumerate(W.initial_fields(prm)):
        g.set_field(v, f)
    return g


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps of the fixed-dt paths")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="beta_alpha_192")
    ap.add_argument("--horizon-steps", type=int, default=1000, help="horizon t_end = dt x this")
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--controller", choices=["learned", "plain"], default="learned")
    ap.add_argument("--weights", default=CONTROLLER)
    ap.add_argument("--precision", type=int, choices=[1, 2], default=2,
                    help="mixed-precision mode products: 1 BF16 split-3, 2 FP16 split-2")
    ap.add_argument("--learned-safety", type=float, default=None,
                    help="margin of the learned proposal below the error model's largest step (library default)")
    ap.add_argument("--quick", action="store_true", help="skip the accuracy runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one plain device copy of this size first (HBM counter calibration)")
    ap.add_argument("--only", choices=["all", "fp64", "mixed-fixed", "headline"], default="all",
                    help="run one part (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: libetd runs one subdomain on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pyetd
    from pyetd import controller as CT
    from pyetd import workloads as W

    prm = getattr(W, args.workload)()
    n, nac, nch = prm["n"], prm["nac"], prm["nch"]
    V, N = nac + nch, n - 2 * prm["g"]
    dt = prm["dt"]
    t_end = dt * args.horizon_steps
    out = {"metric": METRIC, "unit": "sim-time/s", "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "data": "synthetic (seeded noise initial fields, seeded cubic-rotation eigenstrains)"}

    # ---- plain FP64 drop-in path, fixed dt
    fp = None
    if args.only in ("all", "fp64"):
        g = fresh(pyetd, W, prm)
        if args.calib_copy_mib:
            g.calib_copy(args.calib_copy_mib << 20)  # plain copy kernel, exactly this many bytes each way
        if args.warmup:
            g.time_steps(args.warmup, with_diag=True)
        ms, launches = g.time_steps(args.steps, with_diag=True)
        g.close()
        K = args.steps
        mp_flops = 9 * V * 2.0 * n ** 4  # 3 forward (2V fields) + 3 backward (V) mode products per step
        fp = {"sim_time_per_s": dt * K / (ms[0] / 1e3), "steps": K, "warmup": args.warmup, "ms_per_step": ms[0] / K,
              "grid_point_variant_steps_per_s": N ** 3 * V * K / (ms[0] / 1e3),
              "breakdown_ms_per_step": {"mode_products": ms[1] / K, "elastic_fft": ms[2] / K, "fft_passes": ms[7] / K,
                                        "nonlinear": ms[3] / K, "phi_scaling": ms[4] / K, "energy": ms[5] / K,
                                        "halo": ms[6] / K},
              "roofline": {"kernel": "k_modeprod_dl<13,4,2,2> (FP64 MFMA 4x4x4 mode products)", "bound": "mfma",
                           "achieved": mp_flops * K / (ms[1] / 1e3) / 1e12, "peak": PEAK_FP64_MFMA / 1e12,
                           "unit": "TFLOP/s", "frac": mp_flops * K / (ms[1] / 1e3) / PEAK_FP64_MFMA,
                           "traffic": measured_traffic("k_modeprod_dl_192cube_13fields_hbm_counters.md")}}

    # ---- mixed precision, fixed dt: per-kernel timing of the headline path's kernels
    mx = None
    if args.only in ("all", "mixed-fixed"):
        q = dict(prm, precision=args.precision)
        g = fresh(pyetd, W, q)
        if args.calib_copy_mib and args.only == "mixed-fixed":
            g.calib_copy(args.calib_copy_mib << 20)
        if args.warmup:
            g.time_steps(args.warmup, with_diag=True)
        ms, launches = g.time_steps(args.steps, with_diag=True)
        g.close()
        K = args.steps
        mx = {"ms_per_step": ms[0] / K, "sim_time_per_s_fixed_dt": dt * K / (ms[0] / 1e3),
              "grid_point_variant_steps_per_s": N ** 3 * V * K / (ms[0] / 1e3),
              "breakdown_ms_per_step": {"mode_products": ms[1] / K, "elastic_fft": ms[2] / K, "fft_passes": ms[7] / K,
                                        "nonlinear": ms[3] / K, "phi_scaling": ms[4] / K, "energy": ms[5] / K,
                                        "halo": ms[6] / K}}
        # the two largest kernels of the mixed step: FFT axis passes and split mode products
        fft_launch_ms = ms[7] / max(1, launches[6]) / 3.0  # 3 axis passes per fft3d group
        npair = (nac + 1) // 2
        fft_bytes = 2 * npair * N ** 3 * 8.0  # read + write one float2 per point and packed field
        mp_launch_ms = ms[1] / max(1, launches[0])
        mp_fields_avg = (3 * 2 * V + 3 * V) / 6.0
        mp_bytes = 2 * 4.0 * n ** 3 * mp_fields_avg  # read + write FP32 per field
        nterm = 6 if args.precision == 1 else 3  # BF16 split-3: six products; FP16 split-2: three
        mp_mfma = nterm * 2.0 * n ** 4 * mp_fields_avg
        r_fft = {"kernel": "k_fft16<12> (own two-stage register FFT pass, FP32 in HBM)", "bound": "hbm",
                 "achieved": fft_bytes / (fft_launch_ms / 1e3) / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                 "frac": fft_bytes / (fft_launch_ms / 1e3) / PEAK_HBM,
                 "traffic": measured_traffic("k_fft16_f32_192cube_hbm_counters.md"), "launch_ms": fft_launch_ms,
                 "ms_per_step": ms[7] / K}
        r_mp = {"kernel": ("k_modeprod_split_p (split-BF16 x3 MFMA, FP32 accumulate)" if args.precision == 1 else
                           "k_modeprod_f16_r3<13,1> (split-FP16 x2 MFMA, three-slot field ring, FP32 accumulate)"), "bound": "hbm",
                "achieved": mp_bytes / (mp_launch_ms / 1e3) / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": mp_bytes / (mp_launch_ms / 1e3) / PEAK_HBM,
                "traffic": measured_traffic("k_modeprod_split_192cube_hbm_counters.md" if args.precision == 1 else
                                            "k_modeprod_f16_192cube_hbm_counters.md"), "launch_ms": mp_launch_ms,
                "ms_per_step": ms[1] / K, "mfma_tflops": mp_mfma / (mp_launch_ms / 1e3) / 1e12,
                "mfma_frac_16bit": mp_mfma / (mp_launch_ms / 1e3) / PEAK_BF16_MFMA}
        mx["rooflines"] = [r_fft, r_mp]
        out["roofline"] = dict(r_fft if r_fft["ms_per_step"] >= r_mp["ms_per_step"] else r_mp)

    # ---- headline: mixed precision + learned dt over the horizon
    if args.only in ("all", "headline"):
        q = dict(prm, precision=args.precision)
        g = fresh(pyetd, W, q)
        mlp = None
        ctrl = args.controller
        if ctrl == "learned":
            if not os.path.exists(args.weights):
                print(f"bench.py: controller weights {args.weights} missing; run "
                      "python -m pyetd.train_controller", file=sys.stderr)
                sys.exit(1)
            mlp = CT.load(args.weights)
            g.set_controller(mlp)
        kw = dict(t_end=t_end, tol=args.tol, dt0=dt, dt_min=1e-5, dt_max=1.0, controller=ctrl)
        if args.learned_safety is not None:
            kw["learned_safety"] = args.learned_safety
        g.run_adaptive(**dict(kw, t_end=2 * dt))  # warmup: first launches, scratch allocation
        for v, f in enumerate(W.initial_fields(q)):
            g.set_field(v, f)
        g.time = 0.0
        t0 = time.perf_counter()
        st, log = g.run_adaptive(log_cap=100000, **kw)
        wall = time.perf_counter() - t0
        E = log[:, 3]
        out.update({"value": t_end / wall, "steps": st["accepted"], "warmup": 1,
                    "ms_per_step": wall * 1e3 / st["accepted"], "dtype": "f32",
                    "adaptive": dict(st, wall_s=wall, energy_monotone=bool((np.diff(E) <= 1e-7 * np.abs(E[:-1])).all()),
                                     dt_first=float(log[0, 1]), dt_largest=float(log[:, 1].max()),
                                     options={k: getattr(pyetd.adapt_opts(**kw), k)
                                              for k in ("tol", "safety", "learned_safety", "grow", "shrink",
                                                        "extrapolate", "energy_test", "energy_elastic",
                                                        "energy_kind")})})
        headline_fields = [g.get_field(v) for v in range(V)]
        g.close()
        if mlp is not None:
            mlp.close()
            # the same horizon with the plain controller: what the network gains
            g = fresh(pyetd, W, q)
            kw_p = dict(kw, controller="plain")
            g.run_adaptive(**dict(kw_p, t_end=2 * dt))
            for v, f in enumerate(W.initial_fields(q)):
                g.set_field(v, f)
            g.time = 0.0
            t0 = time.perf_counter()
            st_p, _ = g.run_adaptive(**kw_p)
            wall_p = time.perf_counter() - t0
            g.close()
            out["adaptive"]["plain_controller"] = dict(st_p, wall_s=wall_p, sim_time_per_s=t_end / wall_p)
            out["adaptive"]["learned_gain_step_evals"] = st_p["step_evals"] / st["step_evals"]
            out["adaptive"]["learned_gain_wall"] = wall_p / wall
        out["config"] = {"workload": f"{prm['name']}: {N}^3 interior (n={n}), {nac} AC + {nch} CH fields, cubic "
                                     f"elasticity, ETD1, horizon t = {t_end:g} ({args.horizon_steps} steps of dt={dt})",
                         "grid": N, "extended_edge": n, "fields": V, "t_end": t_end,
                         "precision": ("mixed: FP32 fields, " + ("BF16 split-3" if args.precision == 1 else
                                                                  "FP16 split-2 (power-of-two scaled)") +
                                       " MFMA mode products with FP32 accumulation, FP64 energy and acceptance"),
                         "dt_control": f"adaptive, step doubling, tol {args.tol:g}, {ctrl} controller"}
        if fp is not None:
            out["speedup_vs_plain_fp64"] = out["value"] / fp["sim_time_per_s"]
        # ---- fixed-error check against the plain FP64 path, and the mixed fixed-dt report
        if not args.quick:
            a = fresh(pyetd, W, prm)
            b = fresh(pyetd, W, dict(prm, precision=args.precision))
            e_init = float(a.diag()[0][0, 3])
            rows = []
            for k in range(0, args.horizon_steps, 100):
                m = min(100, args.horizon_steps - k)
                a.step(m)
                b.step(m)
                sa, _ = a.diag()
                sb, _ = b.diag()
                deta = max(float(np.abs(a.get_field(v) - b.get_field(v)).max()) for v in range(nac))
                dc = max([float(np.abs(a.get_field(v) - b.get_field(v)).max()) for v in range(nac, V)] or [0.0])
                rows.append({"step": k + m, "max_deta": deta, "max_dc": dc,
                             "rel_energy": abs(sb[0, 3] - sa[0, 3]) / abs(sa[0, 3])})
            ref = [a.get_field(v) for v in range(V)]
            sa, _ = a.diag()
            a.close()
            b.close()
            g2 = fresh(pyetd, W, dict(prm, precision=args.precision))
            for v, f in enumerate(headline_fields):
                g2.set_field(v, f)
            sh, _ = g2.diag()
            g2.close()
            out["accuracy"] = {
                "reference": f"plain FP64 path, {args.horizon_steps} fixed steps",
                "max_deta": max(float(np.abs(headline_fields[v] - ref[v]).max()) for v in range(nac)),
                "max_dc": max([float(np.abs(headline_fields[v] - ref[v]).max()) for v in range(nac, V)] or [0.0]),
                "rel_energy": abs(sh[0, 3] - sa[0, 3]) / abs(sa[0, 3]),
                "energy_fp64": float(sa[0, 3]), "energy_headline": float(sh[0, 3]), "energy_initial": e_init,
                "max_abs_eta_fp64": max(float(np.abs(ref[v]).max()) for v in range(nac)),
                "mixed_fixed_dt_every_100_steps": rows}

    if fp is not None:
        out["plain_fp64"] = fp
    if mx is not None:
        out["mixed_fixed_dt"] = mx
    if "value" not in out:  # partial runs (profiling)
        src = fp if fp is not None else {"sim_time_per_s": mx["sim_time_per_s_fixed_dt"], "steps": args.steps,
                                         "warmup": args.warmup, "ms_per_step": mx["ms_per_step"]}
        out["value"] = src["sim_time_per_s"]
        out["steps"], out["warmup"], out["ms_per_step"] = src["steps"], src["warmup"], src["ms_per_step"]
        out["dtype"] = "f64" if fp is not None else "f32"
        out["config"] = {"workload": f"{prm['name']} fixed dt (partial run --only {args.only})"}
        if "roofline" not in out and fp is not None:
            out["roofline"] = fp["roofline"]

    if not args.no_cpu_baseline and args.only == "all":
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
            "value": dt * args.cpu_steps / cs, "unit": "sim-time/s", "cores": O.threads(), "kind": "port",
            "sample": f"{args.cpu_steps} full fixed-dt step(s) of the same workload (check_soln_new + ETD1 step, "
                      f"{N}^3 x {V} fields) with the FP64 C oracle, OpenMP, {cs:.1f} s"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

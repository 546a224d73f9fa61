#!/usr/bin/env python3
"""Simulated time per wall second of the ScLETD step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload beta_alpha_192]

Workload: 192^3 interior (n = 196) x (12 AC + 1 CH) beta -> alpha fields,
cubic elasticity, ETD1, dt = 0.005, the 1000-step horizon t = 5.  Inputs are
resident in HBM before every timed region.

- value: the headline path, mixed precision (FP32 fields, MFMA mode products
  as FP16 split-2 by default or BF16 split-3 with --precision 1, FP32
  accumulation, FP64 energy) with adaptive steps proposed by the learned
  controller (step doubling, tolerance 1e-4, energy-decrease acceptance).
  One "step" of this line is one adaptive integration of the whole horizon
  t = 0 -> 5 from the seeded initial state: W untimed, then K timed, each
  bracketed by stream synchronisations (the field reset between them is not
  timed); value = K t_end / (sum of the K wall times).
- plain_fp64: the drop-in path (FP64, fixed dt = 0.005): K timed steps after
  W warmup steps, diagnostics included; sim-time/s = dt K / seconds.
- adaptive: the headline run's accepted / rejected steps, and the same
  horizon with the plain controller (safety 0.9, growth 2) and with the plain
  rule retuned to the learned controller's margin and growth bound
  (safety = learned_safety, growth 20): what the network adds beyond them.
- accuracy: the fixed-error check, max|d eta| and max|d c| of the headline
  state against the plain FP64 path after its 1000 steps (the plain FP64 GPU
  path equals the FP64 oracle within 1e-10 at this size,
  tests/test_gpu_configs.py), the relative energy error, monotone energy; and
  the mixed-precision fixed-dt report every 100 steps.
- nucleated: the same comparison from a nucleated state (one seeded nucleus
  per 16^3 cells), where the growth of the nuclei, not a decaying field,
  limits dt; horizon t = 0.5.
- roofline: the dominant kernel of the headline path (the mode products of
  the mixed step, timed at fixed dt with HIP events on the library's stream);
  traffic from the committed rocprofv3 counters.
- cpu_baseline: the FP64 C oracle on this host: the median of 3 full steps of
  the same workload on all host cores, and a 1-core figure on the 64^3 x 4
  configuration (10 steps).  --cpu-full times the 64^3 x 4 configuration in
  full (100 steps, median of 3, 1 core and all cores) and 20 steps of the
  192^3 workload instead.
--equal-error prints instead the nucleated state compared at equal error: the
plain FP64 path at dt / 2^k against a Richardson-extrapolated FP64 reference,
the headline path's error and time on the same state, and the FP64 step (and
speed) that matches the headline's error.
Prints one JSON line.
"""
import argparse
import json
import os
import platform
import statistics
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


def measured_traffic(name, key="traffic_bytes_per_launch_avg"):
    """A figure (HBM bytes per launch, or per field pass) from a committed
    counter summary, or None."""
    path = os.path.join(PROFILES, name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if key + ":" in line:
            return float(line.split(key + ":")[1].strip(" *\n"))
    return None


def traffic_per_launch(name, fields_per_launch):
    """HBM bytes of a mode-product launch of this many fields, from the
    summary's bytes per field pass (a launch of f fields moves f field
    passes)."""
    per_field = measured_traffic(name, "traffic_bytes_per_field_pass")
    return None if per_field is None else per_field * fields_per_launch


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def fresh(pyetd, prm, fields):
    g = pyetd.Etd.from_dict(prm)
    for v, f in enumerate(fields):
        g.set_field(v, f)
    return g


def fixed_dt(pyetd, prm, fields, steps, warmup, calib=0):
    """etd_time_steps: warmup steps, then `steps` steps between HIP events on
    the library's stream (diagnostics included); returns (ms[10], launches[8])."""
    g = fresh(pyetd, prm, fields)
    if calib:
        g.calib_copy(calib)  # plain copy kernel, exactly this many bytes each way
    if warmup:
        g.time_steps(warmup, with_diag=True)
    ms, launches = g.time_steps(steps, with_diag=True)
    g.close()
    return ms, launches


def breakdown(ms, K):
    return {"mode_products": ms[1] / K, "carry_correction": ms[8] / K, "elastic_fft": ms[2] / K,
            "fft_passes": ms[7] / K, "nonlinear": ms[3] / K, "phi_scaling": ms[4] / K, "energy": ms[5] / K,
            "halo": ms[6] / K}


def progress(msg):
    """A line on stderr as each part of the run ends (the JSON line goes to
    stdout at the very end)."""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def horizon_runs(pyetd, prm, fields, kw, count, mlp=None, log_cap=0):
    """`count` adaptive integrations of the horizon from `fields`, each timed
    on its own (set_field and run_adaptive synchronise the stream on both
    sides); returns (wall seconds per run, stats and log of the last run, its
    final fields)."""
    g = pyetd.Etd.from_dict(prm)
    if mlp is not None:
        g.set_controller(mlp)
    walls, st, log = [], None, None
    for _ in range(count):
        for v, f in enumerate(fields):
            g.set_field(v, f)
        g.time = 0.0
        t0 = time.perf_counter()
        st, log = g.run_adaptive(log_cap=log_cap, **kw)
        walls.append(time.perf_counter() - t0)
        progress(f"adaptive integration to t = {kw['t_end']:g}: {st['accepted']} accepted, {walls[-1]:.3f} s")
    out = [g.get_field(v) for v in range(len(fields))]
    g.close()
    return walls, st, log, out


def equal_error(args, pyetd, W, CT, prm, nsteps=100, levels=7):
    """The nucleated state (one seeded nucleus per 16^3 cells, where the
    growth of the nuclei limits dt) over t = 0 -> nsteps dt, compared at equal
    error.  The reference is the plain FP64 path at dt / 2^(levels - 1) and
    dt / 2^(levels - 2), Richardson-extrapolated (ETD1 is first order:
    2 u(h) - u(2h) cancels the O(h) error).  Errors are max |d eta| over the
    interior of the AC fields.  The plain FP64 path runs at dt / 2^k for
    k < levels (timed with HIP events, no diagnostics), the headline path
    (mixed precision + learned controller, tol as given) once; the FP64 step
    that matches the headline's error is found from the finest runs' first-order
    error constant, and the speedup is the headline's sim-time/s over the
    FP64 path's at that step."""
    import numpy as np
    g0, nac, nch = prm["g"], prm["nac"], prm["nch"]
    dt = prm["dt"]
    t_end = dt * nsteps
    nq = dict(prm, init="nuclei", nuclei_per_cell=16.0 ** -3, nuclei_radius=(3.0, 6.0))
    ninit = W.initial_fields(nq)
    inner = (slice(g0, -g0),) * 3

    def fixed(k):
        q = dict(nq, dt=dt / 2 ** k)
        g = fresh(pyetd, q, ninit)
        steps = nsteps * 2 ** k
        g.time_steps(2, with_diag=False)  # warm: first launches and the carried spectrum's buffers
        for v, f in enumerate(ninit):  # restart from the initial state
            g.set_field(v, f)
        ms, _ = g.time_steps(steps, with_diag=False)
        eta = [g.get_field(v)[inner].copy() for v in range(nac)]
        g.close()
        print(f"equal-error: FP64 dt = {q['dt']:.4g}, {steps} steps, {ms[0] / 1e3:.2f} s", file=sys.stderr,
              flush=True)
        return eta, ms[0] / steps

    runs = {k: fixed(k) for k in range(levels)}
    ref = [2.0 * a - b for a, b in zip(runs[levels - 1][0], runs[levels - 2][0])]
    err = lambda eta: max(float(np.abs(a - r).max()) for a, r in zip(eta, ref))
    fixed_rows = [{"dt": dt / 2 ** k, "steps": nsteps * 2 ** k, "ms_per_step": runs[k][1],
                   "sim_time_per_s": (dt / 2 ** k) / (runs[k][1] / 1e3), "max_deta": err(runs[k][0])}
                  for k in range(levels)]
    # the headline path on the same state
    mlp = CT.load(args.weights) if args.controller == "learned" else None
    kw = dict(t_end=t_end, tol=args.tol, dt0=dt, dt_min=1e-5, dt_max=1.0, controller=args.controller)
    horizon_runs(pyetd, dict(nq, precision=args.precision), ninit, kw, 1, mlp)  # warm
    walls, st, log, hf = horizon_runs(pyetd, dict(nq, precision=args.precision), ninit, kw, 1, mlp, log_cap=100000)
    if mlp is not None:
        mlp.close()
    e_head = err([hf[v][inner] for v in range(nac)])
    head = {"precision": args.precision, "controller": args.controller, "tol": args.tol, "wall_s": walls[0],
            "sim_time_per_s": t_end / walls[0], "accepted": st["accepted"],
            "rejected": st["rejected_err"] + st["rejected_energy"], "max_deta": e_head,
            "energy_monotone": monotone(log[:, 3])}
    # first-order error constant from the finest run that is not part of the reference
    kc = levels - 3
    c1 = fixed_rows[kc]["max_deta"] / fixed_rows[kc]["dt"]
    dt_eq = e_head / c1
    ms_fp64 = statistics.median(r["ms_per_step"] for r in fixed_rows)
    fp_eq = dt_eq / (ms_fp64 / 1e3)
    return {"metric": "sim-time/s at equal error, nucleated 192^3 x 13", "unit": "sim-time/s",
            "workload": f"{prm['name']} from {round((prm['n'] - 2 * g0) ** 3 / 16 ** 3)} seeded nuclei "
                        f"(radius 3-6 cells), t = 0 -> {t_end:g}",
            "reference": f"plain FP64 path at dt / {2 ** (levels - 1)} and dt / {2 ** (levels - 2)}, "
                         "Richardson-extrapolated; errors max |d eta| over the interior",
            "reference_self_check_max_deta": err(runs[levels - 1][0]),
            "fixed_dt_fp64": fixed_rows, "headline": head,
            "fp64_first_order_constant": c1, "fp64_dt_at_headline_error": dt_eq,
            "fp64_sim_time_per_s_at_headline_error": fp_eq,
            "speedup_at_equal_error": head["sim_time_per_s"] / fp_eq}


def monotone(E):
    import numpy as np
    E = np.asarray(E)
    return bool((np.diff(E) <= 1e-10 * np.abs(E).max()).all()) if len(E) > 1 else True


def controller_comparison(pyetd, prm, fields, kw, t_end, repeats):
    """Sim-time/s of the plain rule and of the plain rule retuned to the
    learned controller's margin and growth bound, median of `repeats` runs."""
    out = {}
    lsafety = pyetd.adapt_opts(**kw).learned_safety
    for name, extra in (("plain_controller", dict(controller="plain")),
                        ("tuned_plain_controller", dict(controller="plain", safety=lsafety, grow=20.0))):
        w, st, log, _ = horizon_runs(pyetd, prm, fields, dict(kw, **extra), repeats, log_cap=100000)
        med = statistics.median(w)
        out[name] = dict(st, wall_s_median=med, sim_time_per_s=t_end / med, energy_monotone=monotone(log[:, 3]),
                         options={k: v for k, v in extra.items() if k != "controller"})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps (fixed-dt paths) and timed horizon runs")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="beta_alpha_192")
    ap.add_argument("--horizon-steps", type=int, default=1000, help="horizon t_end = dt x this")
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--controller", choices=["learned", "plain"], default="learned")
    ap.add_argument("--weights", default=CONTROLLER)
    ap.add_argument("--precision", type=int, choices=[1, 2], default=2,
                    help="mixed-precision mode products: 1 BF16 split-3, 2 FP16 split-2")
    ap.add_argument("--learned-safety", type=float, default=None,
                    help="margin of the learned proposal below the error model's largest step (library default)")
    ap.add_argument("--repeats", type=int, default=3, help="runs per controller in the headline comparisons")
    ap.add_argument("--nuc-repeats", type=int, default=1, help="runs per controller on the nucleated state")
    ap.add_argument("--quick", action="store_true", help="skip the accuracy runs and the nucleated line")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-full", action="store_true",
                    help="time the CPU oracle in full: 100 steps of the 64^3 x 4 configuration (median of 3, "
                         "1 core and all cores) and 20 steps of the 192^3 workload (several minutes)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one plain device copy of this size first (HBM counter calibration)")
    ap.add_argument("--only", choices=["all", "fp64", "mixed-fixed", "headline"], default="all",
                    help="run one part (for profiling)")
    ap.add_argument("--equal-error", action="store_true",
                    help="instead of the bench line: the nucleated state at equal error (about 3 minutes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: libetd runs one subdomain on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pyetd
    from pyetd import accuracy as AC
    from pyetd import controller as CT
    from pyetd import workloads as W

    prm = getattr(W, args.workload)()
    n, nac, nch = prm["n"], prm["nac"], prm["nch"]
    V, N = nac + nch, n - 2 * prm["g"]
    dt = prm["dt"]
    t_end = dt * args.horizon_steps
    K, Wm = args.steps, args.warmup
    if args.equal_error:
        print(json.dumps(equal_error(args, pyetd, W, CT, prm)))
        return
    init = W.initial_fields(prm)
    out = {"metric": METRIC, "unit": "sim-time/s", "n_gpus": 1, "steps": K, "warmup": Wm, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "data": "synthetic (seeded noise initial fields, seeded cubic-rotation eigenstrains)"}

    # ---- plain FP64 drop-in path, fixed dt
    fp = None
    if args.only in ("all", "fp64"):
        ms, launches = fixed_dt(pyetd, prm, init, K, Wm, args.calib_copy_mib << 20)
        progress(f"plain FP64 fixed dt: {ms[0] / K:.3f} ms per step")
        # field passes the mode products made (3 axes x the fields of each
        # transform: 2V forward and V backward, or nch + V forward when the
        # AC spectrum is carried), counted by the library
        mp_flops = ms[9] * 2.0 * n ** 4
        mp_launch_ms = ms[1] / max(1, launches[0])
        fp = {"sim_time_per_s": dt * K / (ms[0] / 1e3), "steps": K, "warmup": Wm, "ms_per_step": ms[0] / K,
              "grid_point_variant_steps_per_s": N ** 3 * V * K / (ms[0] / 1e3), "breakdown_ms_per_step": breakdown(ms, K),
              "roofline": {"kernel": "k_modeprod_dl<13,4,2,2> (FP64 MFMA 4x4x4 mode products)", "bound": "mfma",
                           "achieved": mp_flops / (ms[1] / 1e3) / 1e12, "peak": PEAK_FP64_MFMA / 1e12,
                           "unit": "TFLOP/s", "frac": mp_flops / (ms[1] / 1e3) / PEAK_FP64_MFMA,
                           "field_passes_per_step": ms[9] / K,
                           "launch_ms_avg": mp_launch_ms, "launches": launches[0],
                           "traffic": traffic_per_launch("k_modeprod_dl_192cube_13fields_hbm_counters.md",
                                                         ms[9] / max(1, launches[0]))}}

    # ---- mixed precision, fixed dt: per-kernel timing of the headline path's kernels
    mx = None
    if args.only in ("all", "mixed-fixed"):
        q = dict(prm, precision=args.precision)
        ms, launches = fixed_dt(pyetd, q, init, K, Wm, (args.calib_copy_mib << 20) if args.only == "mixed-fixed" else 0)
        progress(f"mixed fixed dt: {ms[0] / K:.3f} ms per step")
        mx = {"ms_per_step": ms[0] / K, "sim_time_per_s_fixed_dt": dt * K / (ms[0] / 1e3),
              "grid_point_variant_steps_per_s": N ** 3 * V * K / (ms[0] / 1e3), "breakdown_ms_per_step": breakdown(ms, K)}
        fft_launch_ms = ms[7] / max(1, launches[6]) / 3.0  # 3 axis passes per fft3d group
        npair = (nac + 1) // 2
        fft_bytes = 2 * npair * N ** 3 * 8.0  # read + write one float2 per point and packed field
        mp_launch_ms = ms[1] / max(1, launches[0])
        mp_fields_avg = ms[9] / max(1, launches[0])  # fields per launch (a launch is one axis pass)
        mp_bytes = 2 * 4.0 * n ** 3 * mp_fields_avg  # read + write FP32 per field
        nterm = 6 if args.precision == 1 else 3  # BF16 split-3: six products; FP16 split-2: three
        mp_mfma = nterm * 2.0 * n ** 4 * mp_fields_avg
        r_fft = {"kernel": "k_fft16<12> (own two-stage register FFT pass, FP32 in HBM)", "bound": "hbm",
                 "achieved": fft_bytes / (fft_launch_ms / 1e3) / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                 "frac": fft_bytes / (fft_launch_ms / 1e3) / PEAK_HBM,
                 "traffic": measured_traffic("k_fft16_f32_192cube_hbm_counters.md"), "launch_ms": fft_launch_ms,
                 "ms_per_step": ms[7] / K}
        r_mp = {"kernel": ("k_modeprod_split_p (split-BF16 x3 MFMA, FP32 accumulate)" if args.precision == 1 else
                           "k_modeprod_f16_r3<13,2> (split-FP16 x2 MFMA, contraction over the slowest axis, "
                           "FP32 accumulate)"),
                "bound": "hbm", "achieved": mp_bytes / (mp_launch_ms / 1e3) / 1e9, "peak": PEAK_HBM / 1e9,
                "unit": "GB/s", "frac": mp_bytes / (mp_launch_ms / 1e3) / PEAK_HBM,
                "traffic": traffic_per_launch("k_modeprod_split_192cube_hbm_counters.md" if args.precision == 1 else
                                              "k_modeprod_f16_192cube_hbm_counters.md", mp_fields_avg),
                "launch_ms": mp_launch_ms,
                "launches": launches[0], "fields_per_launch": mp_fields_avg, "ms_per_step": ms[1] / K,
                "mfma_tflops": mp_mfma / (mp_launch_ms / 1e3) / 1e12,
                "mfma_frac_16bit": mp_mfma / (mp_launch_ms / 1e3) / PEAK_BF16_MFMA}
        mx["rooflines"] = [r_fft, r_mp]
        out["roofline"] = dict(r_fft if r_fft["ms_per_step"] >= r_mp["ms_per_step"] else r_mp)

    # ---- headline: mixed precision + learned dt over the horizon
    if args.only in ("all", "headline"):
        q = dict(prm, precision=args.precision)
        mlp = None
        ctrl = args.controller
        if ctrl == "learned":
            if not os.path.exists(args.weights):
                print(f"bench.py: controller weights {args.weights} missing; run "
                      "python -m pyetd.train_controller", file=sys.stderr)
                sys.exit(1)
            mlp = CT.load(args.weights)
        kw = dict(t_end=t_end, tol=args.tol, dt0=dt, dt_min=1e-5, dt_max=1.0, controller=ctrl)
        if args.learned_safety is not None:
            kw["learned_safety"] = args.learned_safety
        if Wm:
            horizon_runs(pyetd, q, init, kw, Wm, mlp)
        walls, st, log, headline_fields = horizon_runs(pyetd, q, init, kw, K, mlp, log_cap=100000)
        wall = sum(walls)
        E = log[:, 3]
        out.update({"value": K * t_end / wall, "ms_per_step": wall * 1e3 / K, "dtype": "f32",
                    "step_definition": f"one adaptive integration of the horizon t = 0 -> {t_end:g} from the seeded "
                                       f"initial fields ({st['accepted']} accepted steps)",
                    "adaptive": dict(st, accepted_steps=st["accepted"], wall_s_per_horizon=walls,
                                     energy_monotone=monotone(E), dt_first=float(log[0, 1]),
                                     dt_largest=float(log[:, 1].max()),
                                     options={k: getattr(pyetd.adapt_opts(**kw), k)
                                              for k in ("tol", "safety", "learned_safety", "grow", "shrink",
                                                        "extrapolate", "energy_test", "energy_elastic", "learned_online",
                                                        "energy_kind")})})
        if mlp is not None:
            cmp_ = controller_comparison(pyetd, q, init, kw, t_end, args.repeats)
            out["adaptive"].update(cmp_)
            med = statistics.median(walls)
            for name in cmp_:
                out["adaptive"][f"learned_gain_vs_{name}"] = cmp_[name]["wall_s_median"] / med
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
            progress(f"accuracy: {args.horizon_steps} fixed steps, plain FP64 and mixed")
            rows, ref, sa = AC.compare(prm, dict(prm, precision=args.precision), args.horizon_steps, every=100,
                                       fields=init, log=None, keep_ref=True)
            g2 = fresh(pyetd, dict(prm, precision=args.precision), headline_fields)
            sh, _ = g2.diag()
            g2.close()
            g3 = fresh(pyetd, prm, init)
            e_init = float(g3.diag()[0][0, 3])
            g3.close()
            out["accuracy"] = {
                "reference": f"plain FP64 path, {args.horizon_steps} fixed steps",
                "max_deta": max(float(np.abs(headline_fields[v] - ref[v]).max()) for v in range(nac)),
                "max_dc": max([float(np.abs(headline_fields[v] - ref[v]).max()) for v in range(nac, V)] or [0.0]),
                "rel_energy": abs(sh[0, 3] - sa[0, 3]) / abs(sa[0, 3]),
                "energy_fp64": float(sa[0, 3]), "energy_headline": float(sh[0, 3]),
                "energy_initial": e_init,
                "max_abs_eta_fp64": max(float(np.abs(ref[v]).max()) for v in range(nac)),
                "mixed_fixed_dt_every_100_steps": rows}
            del ref
            # ---- the nucleated state: dt limited by the growth of the nuclei
            nq = dict(prm, init="nuclei", nuclei_per_cell=16.0 ** -3, nuclei_radius=(3.0, 6.0))
            ninit = W.initial_fields(nq)
            nsteps = 100
            nt = dt * nsteps
            nkw = dict(kw, t_end=nt)
            nw, nst, nlog, nfields = horizon_runs(pyetd, dict(nq, precision=args.precision), ninit, nkw,
                                                  args.nuc_repeats, mlp, log_cap=100000)
            progress(f"nucleated: {nsteps} fixed steps, plain FP64 and mixed")
            nrows, nref, nsa = AC.compare(nq, dict(nq, precision=args.precision), nsteps, every=nsteps,
                                          fields=ninit, log=None, keep_ref=True)
            nmed = statistics.median(nw)
            nuc = {"workload": f"{prm['name']} from {round(N ** 3 / 16 ** 3)} seeded nuclei (radius 3-6 cells), "
                               f"horizon t = {nt:g} ({nsteps} steps of dt={dt})",
                   "sim_time_per_s": nt / nmed, "wall_s_median": nmed, "adaptive": nst,
                   "energy_monotone": monotone(nlog[:, 3]),
                   "max_deta_vs_fp64_fixed_dt": max(float(np.abs(nfields[v] - nref[v]).max()) for v in range(nac)),
                   "max_dc_vs_fp64_fixed_dt": max([float(np.abs(nfields[v] - nref[v]).max())
                                                   for v in range(nac, V)] or [0.0]),
                   "mixed_fixed_dt_vs_fp64": nrows[-1]}
            if fp is not None:
                nuc["speedup_vs_plain_fp64"] = nuc["sim_time_per_s"] / fp["sim_time_per_s"]
            if mlp is not None:
                ncmp = controller_comparison(pyetd, dict(nq, precision=args.precision), ninit, nkw, nt,
                                             args.nuc_repeats)
                nuc.update(ncmp)
                for name in ncmp:
                    nuc[f"learned_gain_vs_{name}"] = ncmp[name]["wall_s_median"] / nmed
            out["nucleated"] = nuc
        if mlp is not None:
            mlp.close()

    if fp is not None:
        out["plain_fp64"] = fp
    if mx is not None:
        out["mixed_fixed_dt"] = mx
    if "value" not in out:  # partial runs (profiling)
        src = fp if fp is not None else {"sim_time_per_s": mx["sim_time_per_s_fixed_dt"], "ms_per_step": mx["ms_per_step"]}
        out["value"] = src["sim_time_per_s"]
        out["ms_per_step"] = src["ms_per_step"]
        out["dtype"] = "f64" if fp is not None else "f32"
        out["config"] = {"workload": f"{prm['name']} fixed dt (partial run --only {args.only})"}
        if "roofline" not in out and fp is not None:
            out["roofline"] = fp["roofline"]

    if not args.no_cpu_baseline and args.only == "all":
        out["cpu_baseline"] = cpu_baseline(W, prm, init, args.cpu_full)
    print(json.dumps(out), flush=True)


def cpu_baseline(W, prm, init, full):
    """The FP64 C oracle (the reference's algorithm restated in C, OpenMP) on
    this host: the same workload on all cores (median of 3 single steps, or 20
    steps with --cpu-full) and the 64^3 x 4 configuration on 1 core (10
    steps, or in full: 100 steps, median of 3, 1 core and all cores)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    all_cores = O.threads()
    dt = prm["dt"]

    def timed(p, fields, steps, threads):
        O.set_threads(threads)
        o = O.Oracle(p)
        for v, f in enumerate(fields):
            o.set_field(v, f)
        t0 = time.perf_counter()
        for _ in range(steps):
            o.diag()
            o.step()
        s = time.perf_counter() - t0
        del o
        print(f"cpu baseline: {steps} steps of n = {p['n']} on {threads} thread(s): {s:.2f} s", file=sys.stderr,
              flush=True)
        return s

    big_steps = 20 if full else 1
    big = [timed(prm, init, big_steps, all_cores) for _ in range(1 if full else 3)]
    c1 = W.beta_omega_64()
    c1_init = W.initial_fields(c1)
    c1_steps = 100 if full else 10
    reps = 3 if full else 1
    c1_one = statistics.median(timed(c1, c1_init, c1_steps, 1) for _ in range(reps))
    res = {"value": dt * big_steps / statistics.median(big), "unit": "sim-time/s", "cores": all_cores, "kind": "port",
           "cpu_model": cpu_model(),
           "sample": (f"{big_steps} full fixed-dt step(s) (check_soln_new + ETD1 step) of the same workload, "
                      f"{prm['n'] - 4}^3 x {prm['nac'] + prm['nch']} fields, FP64 C oracle, OpenMP on {all_cores} "
                      f"threads, median of {len(big)}: {statistics.median(big):.2f} s"),
           "single_core_64cube": {"value": c1["dt"] * c1_steps / c1_one, "unit": "sim-time/s", "cores": 1,
                                  "sample": f"{c1_steps} steps of the 64^3 x 4 configuration on 1 thread, median of "
                                            f"{reps}: {c1_one:.2f} s",
                                  "grid_point_variant_steps_per_s": 64 ** 3 * 4 * c1_steps / c1_one}}
    if full:
        c1_all = statistics.median(timed(c1, c1_init, c1_steps, all_cores) for _ in range(reps))
        res["all_cores_64cube"] = {"value": c1["dt"] * c1_steps / c1_all, "unit": "sim-time/s", "cores": all_cores,
                                   "sample": f"100 steps, median of 3: {c1_all:.2f} s",
                                   "grid_point_variant_steps_per_s": 64 ** 3 * 4 * c1_steps / c1_all}
    O.set_threads(all_cores)
    return res


if __name__ == "__main__":
    main()

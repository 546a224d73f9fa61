"""The learned-controller configuration: a 256^3 beta -> omega coarsening run
(4 variants, mixed precision, 4096 seeded nuclei) with adaptive ETD, step
doubling and the energy test at tol 1e-4, for the same budget of step
attempts with each controller: the plain rule (safety 0.9, growth 2), the
plain rule retuned to the learned controller's margin and growth bound
(safety = learned_safety, growth 20), the learned controller, and the learned
controller without its online correction.  Reports the rejection rate, the
accepted steps and the total simulated time of each.

    python -m pyetd.coarsening_c5 [--attempts 20000] [--weights controller/mlp.npz]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402
import pyetd  # noqa: E402
from pyetd import controller as CT  # noqa: E402
from pyetd import workloads as W  # noqa: E402


def run(prm, attempts, controller, mlp, tol, chunk=1000, **extra):
    """Adaptive run for a budget of step attempts, in chunks of `chunk`
    attempts (each chunk continues from the last accepted dt; progress goes to
    stderr)."""
    g = pyetd.Etd.from_dict(prm)
    for v, f in enumerate(W.initial_fields(prm)):
        g.set_field(v, f)
    if mlp is not None:
        g.set_controller(mlp)
    tot = dict(accepted=0, rejected_err=0, rejected_energy=0, learned_proposals=0, learned_fallbacks=0,
               step_evals=0)
    logs = []
    dt0 = prm["dt"]
    done = 0
    t0 = time.perf_counter()
    while done < attempts:
        k = min(chunk, attempts - done)
        st, log = g.run_adaptive(log_cap=k, t_end=1e9, tol=tol, dt0=dt0, dt_min=1e-5, dt_max=1.0,
                                 controller=controller, max_attempts=k, **extra)
        for key in tot:
            tot[key] += st[key]
        logs.append(log)
        done += st["accepted"] + st["rejected_err"] + st["rejected_energy"]
        dt0 = st["dt_last"]
        print(f"  {controller}: {done} attempts, t = {st['t']:.4g}, dt = {dt0:.3g}, "
              f"{time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    t_end = g.time
    g.close()
    log = np.concatenate(logs)
    E = log[:, 3]
    att = tot["accepted"] + tot["rejected_err"] + tot["rejected_energy"]
    return dict(tot, attempts=att, rejection_rate=(att - tot["accepted"]) / max(1, att), wall_s=wall,
                sim_time=t_end, sim_time_per_s=t_end / wall,
                energy_monotone=bool((np.diff(E) <= 1e-7 * np.abs(E[:-1])).all()),
                dt_median=float(np.median(log[:, 1])), dt_max_used=float(log[:, 1].max()))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--attempts", type=int, default=20000)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--weights", default=os.path.join(os.path.dirname(HERE), "controller", "mlp.npz"))
    ap.add_argument("--n", type=int, default=260)
    args = ap.parse_args()
    prm = W.beta_omega_256_coarsening()
    if args.n != prm["n"]:
        prm = W.scaled(prm, args.n)
    mlp = CT.load(args.weights)
    out = {"config": f"beta -> omega coarsening from seeded nuclei, {args.n - 4}^3 x 4 variants, mixed precision, "
                      f"tol {args.tol:g}, {args.attempts} step attempts per controller"}
    lsafety = pyetd.adapt_opts().learned_safety
    runs = [("plain", "plain", None, {}),
            ("tuned_plain", "plain", None, dict(safety=lsafety, grow=20.0)),
            ("learned", "learned", mlp, {}),
            ("learned_no_online_correction", "learned", mlp, dict(learned_online=0))]
    for name, ctrl, net, extra in runs:
        out[name] = run(prm, args.attempts, ctrl, net, args.tol, **extra)
        out[name]["options"] = extra
        print(json.dumps({name: out[name]}), file=sys.stderr, flush=True)
    for name in ("plain", "tuned_plain", "learned_no_online_correction"):
        out[f"learned_over_{name}_sim_time"] = out["learned"]["sim_time"] / out[name]["sim_time"]
    mlp.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

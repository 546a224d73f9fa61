"""The batched-ensemble configuration: 64 independent beta -> omega
simulations (4 variants, mixed precision) with interior edges drawn from
{32, 48, 64, 80, 96}, per-instance a2 in [0.1, 0.4] and eigenstrain scale in
[0.01, 0.04], 500 fixed steps each, as one hipGraph (30 steps of every
instance per replay; 30 is a multiple of the felas ring period 3), kernel
variants chosen per size by the autotuner.  Reports the aggregate
throughput and each instance's agreement with its solo run.

    python -m pyetd.ensemble_c4 [--instances 64] [--steps 510] [--check 64]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import pyetd  # noqa: E402
from pyetd import workloads as W  # noqa: E402


def instances(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        N = int(rng.choice([32, 48, 64, 80, 96]))
        prm = W.scaled(W.beta_omega_64(seed=seed + k), N + 4)
        prm["landau"] = (float(rng.uniform(0.1, 0.4)), 3.0, 3.0)
        prm["eps0"] = W.eigenstrains(4, float(rng.uniform(0.01, 0.04)), seed + k)
        prm["precision"] = 2  # mixed: FP16 split-2 mode products
        prm["seed"] = seed + k
        out.append(prm)
    return out


def make(prm, tuning, carry=True):
    g = pyetd.Etd.from_dict(prm, carry=carry)
    g.set_tuning(tuning)
    for v, f in enumerate(W.initial_fields(prm)):
        g.set_field(v, f)
    return g


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--steps", type=int, default=510, help="rounded up to a multiple of --per-launch")
    ap.add_argument("--per-launch", type=int, default=30)
    ap.add_argument("--check", type=int, default=64, help="instances compared with solo runs")
    ap.add_argument("--seed", type=int, default=4000)
    args = ap.parse_args()
    prms = instances(args.instances, args.seed)
    t0 = time.perf_counter()
    tunings = {}
    for prm in prms:
        if prm["n"] not in tunings:
            tunings[prm["n"]] = pyetd.autotune(prm, steps=3)
    tune_s = time.perf_counter() - t0
    members = [make(p, tunings[p["n"]]) for p in prms]
    for g in members:  # first launches outside the capture
        g.step(1)
    for g, p in zip(members, prms):
        for v, f in enumerate(W.initial_fields(p)):
            g.set_field(v, f)
        g.time = 0.0
    launches = -(-args.steps // args.per_launch)
    steps = launches * args.per_launch
    E = pyetd.Ensemble(members, args.per_launch)
    E.run(1)  # warm the graph
    for g, p in zip(members, prms):
        for v, f in enumerate(W.initial_fields(p)):
            g.set_field(v, f)
    t0 = time.perf_counter()
    E.run(launches)
    wall = time.perf_counter() - t0
    E.close()
    # the same instances one by one, eager, for agreement and the serial time
    worst = 0.0
    serial = 0.0
    for i in range(min(args.check, len(prms))):
        # the same history as the ensemble member (warm step, reset, one warm
        # graph launch, reset): the felas ring then holds the same values
        # (and, as the graph does, transforming u every step)
        s = make(prms[i], tunings[prms[i]["n"]], carry=False)
        s.step(1)
        for v, f in enumerate(W.initial_fields(prms[i])):
            s.set_field(v, f)
        s.step(args.per_launch)
        for v, f in enumerate(W.initial_fields(prms[i])):
            s.set_field(v, f)
        t1 = time.perf_counter()
        s.step(steps)
        serial += time.perf_counter() - t1
        for v in range(s.V):
            worst = max(worst, float(np.abs(s.get_field(v) - members[i].get_field(v)).max()))
        s.close()
    gpvs = sum((p["n"] - 4) ** 3 * 4 for p in prms) * steps / wall
    sizes = {}
    for p in prms:
        sizes[p["n"] - 4] = sizes.get(p["n"] - 4, 0) + 1
    print(json.dumps({
        "config": "batched ensemble", "instances": len(prms), "interior_edges": sizes, "steps_each": steps,
        "steps_per_graph_launch": args.per_launch, "graph_launches": launches, "wall_s": wall,
        "aggregate_instance_steps_per_s": len(prms) * steps / wall,
        "aggregate_grid_point_variant_steps_per_s": gpvs,
        "serial_eager_wall_s_checked_instances": serial, "checked_instances": min(args.check, len(prms)),
        "max_abs_diff_vs_solo": worst, "autotune_s": tune_s,
        "tunings": {str(k): v.as_dict() for k, v in tunings.items()}}), flush=True)


if __name__ == "__main__":
    main()

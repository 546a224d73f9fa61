"""Train the learned dt controller on the GPU and save its weights.

    python -m pyetd.train_controller --out energy-stability_amd/controller/mlp.npz [--pairs 100000]

Data: plain adaptive runs (mixed precision, FP16 split-2) of seeded synthetic
workloads at 64^3 - 128^3: beta -> omega with 4 variants, a2 in [0.1, 0.4],
eigenstrain scale in [0.01, 0.04], and beta -> alpha with 12 variants + 1 CH,
each either from seeded nuclei (0.5 - 2 per 16^3 cells, radius 3 - 6: growth,
impingement and coarsening, t_end 10 - 100) or from noise of amplitude 0.01
(decay to the beta state, t_end 5 - 20); tolerances 1e-5 - 3e-4; each run
stops at t_end or after 1000 - 3000 step attempts, so that the pairs come
from many runs.  Every accepted step gives one (features, target) pair
(pyetd.controller.targets_from_log).  Training uses libetd's own kernels
(etd_mlp_train: forward, backward, Adam); 5% of the runs are held out and
their loss is reported."""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import pyetd  # noqa: E402
from pyetd import controller as CT  # noqa: E402
from pyetd import workloads as W  # noqa: E402


def run_configs(seed):
    """Endless seeded stream of (workload dict, adaptive options)."""
    rng = np.random.default_rng(seed)
    k = 0
    while True:
        N = int(rng.choice([64, 96, 128]))
        r = rng.random()
        family = ("alpha-decay" if r < 0.15 else "alpha-nuclei" if r < 0.35 else "decay" if r < 0.5
                  else "nuclei")
        t_end = float(rng.choice([5.0, 10.0, 20.0]))
        if not family.startswith("alpha"):
            prm = W.scaled(W.beta_omega_64(seed=1000 + k), N + 4)
            a2 = float(rng.uniform(0.1, 0.4))
            prm["landau"] = (a2, 3.0, 3.0)
            prm["eps0"] = W.eigenstrains(4, float(rng.uniform(0.01, 0.04)), 1000 + k)
        else:
            prm = W.scaled(W.beta_alpha_192(seed=1000 + k), N + 4)
        if family.endswith("nuclei"):
            prm.update(init="nuclei", nuclei_per_cell=float(rng.uniform(0.5, 2.0)) * 16.0 ** -3,
                       nuclei_radius=(3.0, 6.0))
            t_end = float(rng.choice([10.0, 30.0, 100.0]))
        prm["precision"] = 2
        prm["seed"] = 1000 + k
        opts = dict(t_end=t_end, tol=float(rng.choice([1e-5, 3e-5, 1e-4, 3e-4])),
                    dt0=float(rng.choice([0.002, 0.005, 0.01])), dt_min=1e-5, dt_max=1.0,
                    max_attempts=int(rng.choice([1000, 2000, 3000])))
        yield family, N, prm, opts
        k += 1


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--pairs", type=int, default=100000)
    ap.add_argument("--max-seconds", type=float, default=600.0)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--seed", type=int, default=2024)
    args = ap.parse_args()
    X, Y, runs = [], [], []
    t0 = time.time()
    total = 0
    for family, N, prm, opts in run_configs(args.seed):
        st, log = CT.plain_run(prm, **opts)
        x, y = CT.targets_from_log(log, opts["tol"], opts["dt_min"], opts["dt_max"])
        X.append(x)
        Y.append(y)
        runs.append((family, N, opts["tol"], len(y)))
        total += len(y)
        if len(runs) % 20 == 0:
            print(f"  {len(runs)} runs, {total} pairs, {time.time() - t0:.0f} s", flush=True)
        if total >= args.pairs or time.time() - t0 > args.max_seconds:
            break
    gen_s = time.time() - t0
    nrun = len(runs)
    hold = max(1, nrun // 20)
    Xtr, Ytr = np.concatenate(X[:-hold]), np.concatenate(Y[:-hold])
    Xte, Yte = np.concatenate(X[-hold:]), np.concatenate(Y[-hold:])
    print(f"data: {nrun} runs, {total} pairs in {gen_s:.0f} s; train {len(Ytr)}, held out {len(Yte)}", flush=True)
    m = pyetd.Mlp(seed=args.seed)
    t1 = time.time()
    el = CT.fit(m, Xtr, Ytr, epochs=args.epochs, batch=512, lr=2e-3, seed=args.seed)
    p, fm, fs, ym, ys = m.get()
    yhat = m.forward(((Xte - fm) / fs).astype(np.float32)) * ys + ym
    test_rmse = float(np.sqrt(np.mean((yhat - Yte) ** 2)))
    base_rmse = float(np.sqrt(np.mean((Yte - Ytr.mean()) ** 2)))
    print(f"train: {args.epochs} epochs in {time.time() - t1:.1f} s, loss {el[0]:.4f} -> {el[-1]:.4f}; held-out "
          f"RMSE {test_rmse:.4f} decades (predicting the mean: {base_rmse:.4f})", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    CT.save(args.out, m, meta=dict(pairs=total, runs=nrun, epoch_loss=el, heldout_rmse=test_rmse,
                                   heldout_rmse_mean_predictor=base_rmse, seed=args.seed))
    print(f"saved {args.out}", flush=True)


if __name__ == "__main__":
    main()

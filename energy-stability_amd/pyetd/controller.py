"""Training data, training and storage of the learned dt controller.

Data: plain adaptive runs (etd_run_adaptive with the plain controller) log,
for every accepted step k, the ETD_NFEAT features seen when dt_k was proposed
and the step's error estimate err_k.  The error model's largest step for that
state is dt_best = dt_k sqrt(tol / err_k) (ETD1's local error grows as dt^2),
clipped to [dt_min, dt_max]: what a perfect controller would have proposed.
The target is log10(dt_best / dt_plain), dt_plain being the plain
controller's proposal for the same step (feature 15), so the network learns
a correction to the plain rule, which the adaptive stepper bounds to a decade.  The network is trained on the GPU by libetd's own kernels
(etd_mlp_train) on standardised features and targets."""
import numpy as np

from . import NFEAT, MLP_PARAMS, Etd, Mlp
from . import workloads as W


def targets_from_log(log, tol, dt_min, dt_max):
    """(features, target) pairs of one run's log rows [t, dt, err, E, features]."""
    log = np.asarray(log)
    if log.size == 0:
        return np.zeros((0, NFEAT)), np.zeros(0)
    dt, err = log[:, 1], np.maximum(log[:, 2], 1e-300)
    best = np.clip(dt * np.sqrt(tol / err), dt_min, dt_max)
    feats = log[:, 4:4 + NFEAT].copy()
    return feats, np.log10(best) - feats[:, 15]


def plain_run(prm, t_end, tol, dt0, dt_min, dt_max, log_cap=100000, **kw):
    """One plain adaptive run of workload prm from its seeded initial state."""
    g = Etd.from_dict(prm)
    for v, f in enumerate(W.initial_fields(prm)):
        g.set_field(v, f)
    st, log = g.run_adaptive(log_cap=log_cap, t_end=t_end, tol=tol, dt0=dt0, dt_min=dt_min, dt_max=dt_max,
                             controller="plain", **kw)
    g.close()
    return st, log


def fit(mlp, X, y, epochs=200, batch=256, lr=2e-3, seed=7):
    """Standardises, trains on the GPU, stores the normalisation; returns the
    per-epoch loss (standardised units)."""
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    fm = X.mean(0)
    fs = X.std(0)
    fs[fs < 1e-6] = 1.0
    ym, ys = float(y.mean()), float(max(y.std(), 1e-6))
    p, _, _, _, _ = mlp.get()
    mlp.set(p, fm, fs, ym, ys)
    Xn = ((X - fm) / fs).astype(np.float32)
    yn = ((y - ym) / ys).astype(np.float32)
    return mlp.train(Xn, yn, epochs=epochs, batch=min(batch, len(yn)), lr=lr, seed=seed)


def save(path, mlp, meta=None):
    p, fm, fs, ym, ys = mlp.get()
    np.savez(path, params=p, feat_mean=fm, feat_std=fs, y_mean=np.float32(ym), y_std=np.float32(ys),
             **({} if meta is None else {k: np.asarray(v) for k, v in meta.items()}))


def load(path, device=0):
    z = np.load(path, allow_pickle=False)
    assert z["params"].size == MLP_PARAMS
    m = Mlp(device)
    m.set(z["params"], z["feat_mean"], z["feat_std"], float(z["y_mean"]), float(z["y_std"]))
    return m

"""Synthetic workloads (parameter dictionaries shared by the GPU binding and the
CPU oracle).  Every random choice is seeded.

beta_omega_64   N = 64 (n = 68), 4 variants, FP64, h = 1, dt = 0.01, 100 steps;
                Landau a2 = 0.2, a3 = 3, a4 = 3, kappa (epn2) = 1, L = 1;
                cubic C11 = 134, C12 = 110, C44 = 36; eigenstrains: one seeded
                symmetric tensor (entries in +/-0.03) and its images under the
                cubic rotation group; eta ~ U[-0.01, 0.01].
beta_alpha_192  N = 192 (n = 196), 12 variants + 1 CH concentration
                (mean 0.1, noise +/-0.005), eigenstrain entries +/-0.05,
                dt = 0.005, 1000 steps, ETD1, periodic, FP64.
beta_omega_256_coarsening
                N = 256 (n = 260), 4 variants, mixed precision, dt = 0.01;
                seeded nuclei (init "nuclei"): 1 nucleus per 16^3 cells (4096
                at N = 256), each a sphere of one random variant, radius
                U[3, 6] cells (the critical radius is about 2.5), tanh
                profile of width 1, on eta ~ U[-0.01, 0.01].  The nuclei grow,
                impinge and coarsen; plain noise of any amplitude below the
                barrier (eta = 0.072 for a2 = 0.2, a3 = a4 = 3) decays to the
                beta state instead, as the reference seeds its omega run with
                a sphere (field.c:765, init_field_sphere_center).
The stabiliser KE (not fixed by the configurations) is 3.2 = max f'' of the
single-variant Landau polynomial on [0, 1]; CH uses L = 1, KCI = 2.
"""
import itertools

import numpy as np


def cubic_voigt(c11, c12, c44):
    C = np.zeros((6, 6))
    C[:3, :3] = c12
    for i in range(3):
        C[i, i] = c11
        C[3 + i, 3 + i] = c44
    return C


def cubic_rotations():
    """The 24 proper rotations of the cube as signed permutation matrices, in a
    fixed order (identity first)."""
    mats = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            R = np.zeros((3, 3))
            for i in range(3):
                R[i, perm[i]] = signs[i]
            if np.isclose(np.linalg.det(R), 1.0):
                mats.append(R)
    return mats


def eigenstrains(nvar, amp, seed):
    """nvar distinct images R e R^T of one seeded symmetric base tensor."""
    rng = np.random.default_rng(seed)
    base = rng.uniform(-amp, amp, size=(3, 3))
    base = 0.5 * (base + base.T)
    out = []
    for R in cubic_rotations():
        e = R @ base @ R.T
        if not any(np.allclose(e, x) for x in out):
            out.append(e)
        if len(out) == nvar:
            break
    while len(out) < nvar:  # a symmetric base has at least 12 distinct images
        out.append(out[len(out) % len(out)])
    return np.array(out)


def beta_omega_64(seed=2024):
    nac = 4
    return dict(name="beta_omega_64", n=68, g=2, nac=nac, nch=0, etd2=0, elastic=1, aniso=0, dt=0.01, steps=100,
                epn2=1.0, h=(1.0, 1.0, 1.0), kk=(1.0, 1.0, 1.0), LE=1.0, KE=3.2, landau=(0.2, 3.0, 3.0), LCI=1.0,
                KCI=2.0, elastic_scale=1.0, C=cubic_voigt(134.0, 110.0, 36.0),
                eps0=eigenstrains(nac, 0.03, seed), lam=np.array([np.eye(3)] * nac), seed=seed)


def beta_alpha_192(seed=2024):
    nac = 12
    return dict(name="beta_alpha_192", n=196, g=2, nac=nac, nch=1, etd2=0, elastic=1, aniso=0, dt=0.005, steps=1000,
                epn2=1.0, h=(1.0, 1.0, 1.0), kk=(1.0, 1.0, 1.0), LE=1.0, KE=3.2, landau=(0.2, 3.0, 3.0), LCI=1.0,
                KCI=2.0, elastic_scale=1.0, C=cubic_voigt(134.0, 110.0, 36.0),
                eps0=eigenstrains(nac, 0.05, seed), lam=np.array([np.eye(3)] * nac), seed=seed)


def beta_omega_256_coarsening(seed=2024):
    prm = beta_omega_64(seed)
    prm.update(name="beta_omega_256_coarsening", n=260, precision=2, init="nuclei", nuclei_per_cell=16.0 ** -3,
               nuclei_radius=(3.0, 6.0), steps=20000)
    return prm


def scaled(prm, n, **over):
    """Same physics on a smaller extended edge n (parity runs)."""
    q = dict(prm)
    q["n"] = n
    q.update(over)
    return q


def seed_nuclei(fields, prm, rng):
    """Add seeded spherical nuclei to the interior AC fields (in place): each a
    random variant p, centre uniform on the periodic box, radius uniform in
    prm["nuclei_radius"], profile eta_p = max(eta_p, (1 - tanh((d - r) / w)) / 2)
    with w = 1 and d the periodic distance to the centre."""
    N = fields[0].shape[0]
    count = max(1, int(round(prm.get("nuclei_per_cell", 16.0 ** -3) * N ** 3)))
    rlo, rhi = prm.get("nuclei_radius", (3.0, 6.0))
    w = 1.0
    for _ in range(count):
        c = rng.uniform(0, N, size=3)
        r = rng.uniform(rlo, rhi)
        p = int(rng.integers(prm["nac"]))
        ext = int(np.ceil(r + 4 * w))
        axes = [np.arange(int(np.floor(ci)) - ext, int(np.floor(ci)) + ext + 2) for ci in c]
        dz, dy, dx = np.meshgrid(axes[0] - c[0], axes[1] - c[1], axes[2] - c[2], indexing="ij")
        d = np.sqrt(dz ** 2 + dy ** 2 + dx ** 2)
        prof = 0.5 * (1.0 - np.tanh((d - r) / w))
        ix = np.ix_(axes[0] % N, axes[1] % N, axes[2] % N)
        fields[p][ix] = np.maximum(fields[p][ix], prof)  # [z][y][x]


def initial_fields(prm):
    """Seeded initial state as host arrays (AC: U[-a, a] with a = init_amp,
    default 0.01, plus nuclei when init is "nuclei"; CH: 0.1 +
    U[-0.005, 0.005]); interior drawn, ghosts the periodic images."""
    n, g = prm["n"], prm.get("g", 2)
    N = n - 2 * g
    rng = np.random.default_rng(prm.get("seed", 0) + 17)
    idx = (np.arange(n) - g) % N
    inner = []
    for v in range(prm["nac"] + prm["nch"]):
        if v < prm["nac"]:
            a = prm.get("init_amp", 0.01)
            inner.append(rng.uniform(-a, a, size=(N, N, N)))
        else:
            inner.append(0.1 + rng.uniform(-0.005, 0.005, size=(N, N, N)))
    if prm.get("init") == "nuclei" and prm["nac"] > 0:
        seed_nuclei(inner, prm, rng)
    return [P[np.ix_(idx, idx, idx)] for P in inner]

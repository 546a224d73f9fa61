import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "energy-stability_amd"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


# Landau constants hard-coded in source_beta_to_omega/src/formula.c:126-128
OMEGA_LANDAU = (2376.0 / 2372.0, (7128.0 + 12.0 * 2372.0) / 2372.0, (4752.0 + 12.0 * 2372.0) / 2372.0)


def cubic_voigt(c11, c12, c44):
    C = np.zeros((6, 6))
    C[:3, :3] = c12
    for i in range(3):
        C[i, i] = c11
        C[3 + i, 3 + i] = c44
    return C


def splitmix64_uniform(seed, count):
    """splitmix64 stream -> doubles in [0, 1) via (x >> 11) * 2^-53."""
    out = np.empty(count, dtype=np.float64)
    x = np.uint64(seed)
    mask = np.uint64(0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        for i in range(count):
            x = (x + np.uint64(0x9E3779B97F4A7C15)) & mask
            z = x
            z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & mask
            z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & mask
            z = z ^ (z >> np.uint64(31))
            out[i] = float(z >> np.uint64(11)) * (1.0 / 9007199254740992.0)
    return out


def small_params(n=12, nac=2, nch=0, etd2=0, elastic=0, aniso=0, dt=0.01, seed=1, kk=(1.0, 1.0, 1.0),
                 landau=(0.2, 3.0, 3.0), KE=3.2, epn2=1.0, h=(1.0, 1.0, 1.0)):
    rng = np.random.default_rng(seed)
    eps = rng.uniform(-0.03, 0.03, size=(max(nac, 1), 3, 3))
    eps = 0.5 * (eps + eps.transpose(0, 2, 1))
    lam = np.array([np.diag(kk) for _ in range(max(nac, 1))])
    if aniso:
        lin = rng.uniform(-0.3, 0.3, size=(max(nac, 1), 3, 3))
        lam = lam - lin
    return dict(n=n, g=2, nac=nac, nch=nch, etd2=etd2, elastic=elastic, aniso=aniso, dt=dt, epn2=epn2,
                h=h, kk=kk, LE=1.0, KE=KE, landau=landau, LCI=1.0, KCI=2.0, elastic_scale=1.0,
                C=cubic_voigt(134.0, 110.0, 36.0), eps0=eps[:max(nac, 1)], lam=lam)


def noise_fields(prm, seed=7, amp=0.01, ch_mean=0.1, ch_amp=0.005):
    n = prm["n"]
    rng = np.random.default_rng(seed)
    out = []
    for v in range(prm["nac"]):
        out.append(rng.uniform(-amp, amp, size=(n, n, n)))
    for v in range(prm["nch"]):
        out.append(ch_mean + rng.uniform(-ch_amp, ch_amp, size=(n, n, n)))
    return out


@pytest.fixture
def golden_dir():
    return GOLDEN

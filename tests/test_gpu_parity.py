"""GPU parity: every HIP kernel and the full ETD1 / ETD2 step, called through
the C ABI (pyetd -> libetd.so), against the FP64 CPU oracle on the same seeded
inputs.  The plain FP64 path must agree with the oracle within 1e-10 (fields
are O(1e-2 .. 1)); single kernels within 1e-12 relative to the data scale."""
import sys

import numpy as np
import pytest

import oracle as O
import pyetd
from conftest import OMEGA_LANDAU, noise_fields, small_params, splitmix64_uniform

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pair(prm, fields=None, seed=7):
    """Oracle and GPU contexts with identical parameters and fields."""
    o = O.Oracle(prm)
    Vm, Vim, D = O.eigbasis(prm["n"], 3)
    g = pyetd.Etd(pyetd.params_from_dict(prm), Vm, Vim, D)
    if fields is None:
        fields = noise_fields(prm, seed=seed)
    for v, f in enumerate(fields):
        o.set_field(v, f)
        g.set_field(v, f)
    return o, g


@pytest.mark.parametrize("n", [12, 21, 36, 68])
@pytest.mark.parametrize("rotate", [True, False])
def test_modeprod_matches_oracle(n, rotate):
    say(f"modeprod n={n} rotate={rotate}")
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    u = rng.standard_normal(n ** 3)
    prm = small_params(n=n, nac=1)
    g = pyetd.Etd.from_dict(prm)
    got = g.modeprod(A, u, rotate)
    ref = O.modeprod(A, u, rotate)
    assert np.abs(got - ref).max() < 1e-12 * np.sqrt(n) * np.abs(ref).max()


def test_ghost_fill_bit_exact():
    say("ghost fill")
    prm = small_params(n=14, nac=2, nch=1)
    rng = np.random.default_rng(1)
    f = [rng.standard_normal((14, 14, 14)) for _ in range(3)]
    o, g = pair(prm, f)
    o.transfer()
    g.transfer()
    for v in range(3):
        assert np.array_equal(g.get_field(v), o.get_field(v))


@pytest.mark.parametrize("aniso", [0, 1])
def test_nonlinear_ac_matches_oracle(aniso):
    say(f"nonlinear ac aniso={aniso}")
    prm = small_params(n=16, nac=3, aniso=aniso, h=(0.9, 1.0, 1.2), kk=(1.0, 1.3, 0.7))
    o, g = pair(prm, noise_fields(prm, amp=0.5))
    o.transfer()
    g.transfer()
    for v in range(3):
        ref = o.nonlinear(v)
        got = g.nonlinear(v)
        assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def test_nonlinear_ch_matches_oracle():
    say("nonlinear ch")
    prm = small_params(n=16, nac=2, nch=1, h=(0.8, 1.0, 1.1))
    o, g = pair(prm, noise_fields(prm, amp=0.3, ch_amp=0.3))
    o.transfer()
    g.transfer()
    ref = o.nonlinear(2)
    got = g.nonlinear(2)
    assert np.abs(got - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n", [10, 12, 20, 36, 52, 100, 196])
def test_fft3d_matches_numpy(n):
    """Kernel (d): own FFT (N = 6, 8, 16 on radix passes; 32, 48, 96, 192 on
    the two-stage register kernel, R2 = 2, 3, 6, 12) vs numpy."""
    N = n - 4
    say(f"fft3d N={N}")
    g = pyetd.Etd.from_dict(small_params(n=n, nac=2, elastic=1))
    rng = np.random.default_rng(N)
    x = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    f = g.fft3d(x, 1)
    ref = np.fft.fftn(x)
    assert np.abs(f - ref).max() < 1e-12 * np.abs(ref).max()
    b = g.fft3d(ref, -1)
    assert np.abs(b / N ** 3 - x).max() < 1e-13


@pytest.mark.parametrize("nac,n", [(2, 12), (3, 16), (5, 22), (3, 20), (4, 36), (5, 52)])
def test_elastic_matches_oracle_over_two_evaluations(nac, n):
    """felas on every extended cell, including the ghost cells that inherit
    earlier variants through the shared transfer buffer."""
    say(f"elastic nac={nac} n={n}")
    prm = small_params(n=n, nac=nac, elastic=1)
    o, g = pair(prm, noise_fields(prm, amp=0.4))
    for ev in range(2):
        o.transfer()
        g.transfer()
        o.elastic()
        got = g.elastic()
        for p in range(nac):
            ref = o.felas(p)
            scale = max(1e-3, np.abs(ref).max())
            assert np.abs(got[p] - ref).max() < 1e-12 * scale, (ev, p)
        # perturb the fields so the second evaluation differs
        for v in range(nac):
            f = o.get_field(v) * 1.1 + 0.01
            o.set_field(v, f)
            g.set_field(v, f)


def test_phiscale_matches_oracle():
    say("phiscale")
    prm = small_params(n=12, nac=1, nch=1, kk=(1.0, 1.2, 0.9), h=(1.0, 0.8, 1.1), dt=0.05)
    o, g = pair(prm)
    rng = np.random.default_rng(3)
    n = 12
    for v in range(2):
        uh = rng.standard_normal((n, n, n))  # rotated layout [k'][j'][i']
        nh = rng.standard_normal((n, n, n))
        got = g.phiscale(v, uh, nh).reshape(n, n, n)
        # oracle layout l = j' n^2 + i' n + k'  ->  array [j'][i'][k']
        ou = np.ascontiguousarray(uh.transpose(1, 2, 0)).reshape(-1).copy()
        on = np.ascontiguousarray(nh.transpose(1, 2, 0)).reshape(-1).copy()
        O.lib().orc_phi_update(o.ctx, v, ou.ctypes.data_as(O._dp), on.ctypes.data_as(O._dp))
        ref = ou.reshape(n, n, n).transpose(2, 0, 1)
        assert np.abs(got - ref).max() < 1e-13 * max(1.0, np.abs(ref).max())


CASES = [
    dict(n=12, nac=2, nch=0, etd2=0, elastic=0, aniso=0, steps=6),
    dict(n=12, nac=2, nch=0, etd2=1, elastic=0, aniso=1, steps=6),
    dict(n=16, nac=3, nch=1, etd2=1, elastic=1, aniso=0, steps=5),
    # N = 16: divisible by 16 but without a two-stage FFT instance (R2 = 1):
    # elasticity must take the pack / radix / unpack path
    dict(n=20, nac=3, nch=1, etd2=1, elastic=1, aniso=0, steps=3),
    dict(n=14, nac=4, nch=0, etd2=0, elastic=1, aniso=1, steps=5),
    dict(n=36, nac=4, nch=1, etd2=1, elastic=1, aniso=0, steps=8),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "n{n}_ac{nac}_ch{nch}_etd2{etd2}_el{elastic}_an{aniso}".format(**c))
def test_step_matches_oracle(case):
    say(f"step {case}")
    c = dict(case)
    steps = c.pop("steps")
    prm = small_params(dt=0.02, **c)
    o, g = pair(prm, noise_fields(prm, amp=0.3))
    for s in range(steps):
        o.step()
        g.step()
    worst = 0.0
    for v in range(prm["nac"] + prm["nch"]):
        d = np.abs(g.get_field(v) - o.get_field(v)).max()
        worst = max(worst, d)
    say(f"  max |d| = {worst:.3e}")
    assert worst < STEP_TOL


def test_beta_omega_64cube_elastic_etd2_matches_oracle():
    """The CPU-runnable configuration's shape: N = 64 (n = 68), four variants,
    cubic elasticity, omega Landau constants, ETD2; 3 steps."""
    say("68^3 x 4, elastic, ETD2")
    prm = small_params(n=68, nac=4, etd2=1, elastic=1, dt=0.01, landau=(0.2, 3.0, 3.0), KE=3.2)
    o, g = pair(prm, noise_fields(prm, amp=0.01))
    for s in range(3):
        o.step()
        g.step()
    worst = max(np.abs(g.get_field(v) - o.get_field(v)).max() for v in range(4))
    say(f"  max |d| = {worst:.3e}")
    assert worst < STEP_TOL


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.9, 1.0, 1.2)])
def test_diag_matches_oracle(h):
    """check_soln_new statistics and energy; with non-unit spacing the GPU
    multiplies by 1 / h where the reference divides by h (one rounding)."""
    say(f"diag h={h}")
    prm = small_params(n=14, nac=5, nch=1, landau=OMEGA_LANDAU, aniso=1, h=h)
    o, g = pair(prm, noise_fields(prm, amp=0.5))
    st_o, cp_o = o.diag()
    st_g, cp_g = g.diag()
    assert np.array_equal(st_o[:, :2], st_g[:, :2])
    assert np.abs(st_o[:, 2:] - st_g[:, 2:]).max() < 1e-12 * np.abs(st_o[:, 2:]).max()
    assert abs(cp_o - cp_g) < 1e-12 * abs(cp_o)


@pytest.mark.parametrize("n", [68, 100, 196, 260])
@pytest.mark.parametrize("precision", [0, 2])
def test_diag_matches_oracle_wide(n, precision):
    """The energy kernel at the sizes where a row spans 2 to 5 waves and a
    chunk of planes spans one to two register sub-ranges (rows of n - 5
    cells, chunks of (n - 5) / 8 planes); AC, CH and couple sums."""
    say(f"diag n={n} precision={precision}")
    prm = dict(small_params(n=n, nac=3, nch=1, landau=OMEGA_LANDAU, aniso=1, h=(0.9, 1.0, 1.2)),
               precision=precision)
    f = noise_fields(prm, amp=0.5)
    if precision:  # FP32 state: inputs exactly representable in FP32, sums in FP64
        f = [x.astype(np.float32).astype(np.float64) for x in f]
    o = O.Oracle(dict(prm, precision=0))
    Vm, Vim, D = O.eigbasis(n, 3)
    g = pyetd.Etd(pyetd.params_from_dict(prm), Vm, Vim, D)
    for v, x in enumerate(f):
        o.set_field(v, x)
        g.set_field(v, x)
    st_o, cp_o = o.diag()
    st_g, cp_g = g.diag()
    g.close()
    assert np.array_equal(st_o[:, :2], st_g[:, :2])
    assert np.abs(st_o[:, 2:] - st_g[:, 2:]).max() < 1e-12 * np.abs(st_o[:, 2:]).max()
    assert abs(cp_o - cp_g) < 1e-12 * abs(cp_o)


def test_init_noise_matches_splitmix_reference():
    say("init noise")
    prm = small_params(n=10, nac=1)
    g = pyetd.Etd.from_dict(prm)
    g.init_noise(0, 12345, 0.0, 0.01)
    got = g.get_field(0)
    r = splitmix64_uniform(12345, 6 ** 3).reshape(6, 6, 6)
    P = 0.01 * (2.0 * r - 1.0)
    idx = (np.arange(10) - 2) % 6
    assert np.array_equal(got, P[np.ix_(idx, idx, idx)])

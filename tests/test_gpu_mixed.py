"""GPU parity of the mixed-precision path (FP32 fields, split-BF16 MFMA mode
products with FP32 accumulation, FP64 energy), called through the C ABI on a
context created with precision = ETD_PREC_MIXED, against the FP64 CPU oracle.

Tolerances follow from the formats: FP32 storage rounds at 6e-8 relative; the
split-3 product keeps every term above 2^-24; split-2 keeps 2^-16 (2.5e-5
relative per product, measured on the CPU in the split arithmetic).
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
import pyetd
from conftest import noise_fields, small_params

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def mixed(prm):
    q = dict(prm)
    q["precision"] = 1
    return q


def pair(prm, fields, precision=1):
    o = O.Oracle(prm)
    Vm, Vim, D = O.eigbasis(prm["n"], 3)
    g = pyetd.Etd(pyetd.params_from_dict(dict(prm, precision=precision)), Vm, Vim, D)
    for v, f in enumerate(fields):
        o.set_field(v, f)
        g.set_field(v, f)
    return o, g


@pytest.mark.parametrize("n", [12, 36, 68, 100, 260])
@pytest.mark.parametrize("rotate", [True, False])
def test_split3_modeprod_matches_oracle(n, rotate):
    say(f"split-3 modeprod n={n} rotate={rotate}")
    rng = np.random.default_rng(n + 1)
    Vm, Vim, D = O.eigbasis(n, 3)
    u = rng.standard_normal(n ** 3)
    g = pyetd.Etd.from_dict(mixed(small_params(n=n, nac=1)))
    for A in (Vim, Vm):
        got = g.modeprod(A, u, rotate)
        ref = O.modeprod(A, u, rotate)
        # FP32 input rounding and FP32 accumulation over n terms
        assert np.abs(got - ref).max() < 2e-6 * np.abs(ref).max()


def test_split2_modeprod_matches_oracle():
    say("split-2 modeprod n=68")
    os.environ["ETD_SPLIT"] = "2"
    try:
        n = 68
        rng = np.random.default_rng(5)
        Vm, Vim, D = O.eigbasis(n, 3)
        u = rng.standard_normal(n ** 3)
        g = pyetd.Etd.from_dict(mixed(small_params(n=n, nac=1)))
        got = g.modeprod(Vim, u, True)
    finally:
        del os.environ["ETD_SPLIT"]
    ref = O.modeprod(Vim, u, True)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    say(f"  split-2 rel err {err:.2e}")
    assert err < 2e-4


def test_mixed_fft3d_matches_numpy():
    say("fp32 fft3d N=32")
    n = 36
    g = pyetd.Etd.from_dict(mixed(small_params(n=n, nac=2, elastic=1)))
    rng = np.random.default_rng(11)
    x = rng.standard_normal((32, 32, 32)) + 1j * rng.standard_normal((32, 32, 32))
    f = g.fft3d(x, 1)
    ref = np.fft.fftn(x)
    assert np.abs(f - ref).max() < 2e-6 * np.abs(ref).max()


def test_mixed_nonlinear_and_elastic_match_oracle():
    say("fp32 nonlinear + elastic nac=3 n=16")
    prm = small_params(n=16, nac=3, nch=1, elastic=1, aniso=1, h=(0.9, 1.0, 1.1))
    o, g = pair(prm, noise_fields(prm, amp=0.4, ch_amp=0.3))
    o.transfer()
    g.transfer()
    o.elastic()
    got = g.elastic()
    for p in range(3):
        ref = o.felas(p)
        assert np.abs(got[p] - ref).max() < 1e-5 * max(1e-3, np.abs(ref).max())
    for v in range(4):
        ref = o.nonlinear(v)
        gv = g.nonlinear(v)
        assert np.abs(gv - ref).max() < 1e-5 * max(1.0, np.abs(ref).max()), v


MIXED_CASES = [
    dict(n=16, nac=3, nch=1, etd2=1, elastic=1, aniso=0, steps=5),
    dict(n=20, nac=3, nch=1, etd2=0, elastic=1, aniso=0, steps=4),  # N = 16: no two-stage FFT instance
    dict(n=36, nac=4, nch=1, etd2=0, elastic=1, aniso=1, steps=8),
]


@pytest.mark.parametrize("case", MIXED_CASES, ids=lambda c: "n{n}_ac{nac}_ch{nch}_etd2{etd2}_el{elastic}_an{aniso}".format(**c))
def test_mixed_step_matches_oracle(case):
    say(f"mixed step {case}")
    c = dict(case)
    steps = c.pop("steps")
    prm = small_params(dt=0.02, **c)
    o, g = pair(prm, noise_fields(prm, amp=0.3))
    for _ in range(steps):
        o.step()
        g.step()
    worst = max(np.abs(g.get_field(v) - o.get_field(v)).max() for v in range(prm["nac"] + prm["nch"]))
    say(f"  max |d| = {worst:.3e}")
    # FP32 state: a few 1e-7 per step on O(0.3) fields
    assert worst < 1e-5


def test_mixed_diag_energy_in_fp64():
    say("mixed diag")
    prm = small_params(n=16, nac=3, nch=1, aniso=1)
    fields = noise_fields(prm, amp=0.5)
    fields = [f.astype(np.float32).astype(np.float64) for f in fields]  # exactly representable in FP32
    o, g = pair(prm, fields)
    st_o, cp_o = o.diag()
    st_g, cp_g = g.diag()
    # the inputs are exact in FP32 and the sums run in FP64: FP64 agreement
    assert np.abs(st_o - st_g).max() < 1e-12 * np.abs(st_o).max()
    assert abs(cp_o - cp_g) < 1e-12 * abs(cp_o)



@pytest.mark.parametrize("etd2", [0, 1])
def test_mixed_path_conserves_the_ch_mean(etd2):
    """The CH concentration's interior mean is conserved by the scheme (its
    k = 0 mode has H = 0); the mixed path restores it after every step, so it
    stays at its initial value to FP32 rounding over many steps, as the FP64
    path keeps it to FP64 rounding."""
    say(f"mixed CH mean conservation etd2={etd2}")
    prm = small_params(n=20, nac=2, nch=1, etd2=etd2, elastic=1, dt=0.02)
    f = noise_fields(prm, amp=0.2, ch_amp=0.05)
    g = pyetd.Etd.from_dict(mixed(prm))
    for v, x in enumerate(f):
        g.set_field(v, x)
    gg = prm["g"]
    inner = (slice(gg, -gg),) * 3
    m0 = float(np.mean(g.get_field(2)[inner], dtype=np.float64))
    g.step(60)
    m1 = float(np.mean(g.get_field(2)[inner], dtype=np.float64))
    say(f"  mean {m0:.12f} -> {m1:.12f} (diff {m1 - m0:.2e})")
    assert abs(m1 - m0) < 2e-7 * abs(m0)
    g.close()


@pytest.mark.parametrize("n", [12, 36, 68, 100, 260])
@pytest.mark.parametrize("rotate", [True, False])
def test_f16_split2_modeprod_matches_oracle(n, rotate):
    """FP16 split-2 mode products (ETD_PREC_MIXED_F16): power-of-two scaled V
    and field column chunks, three FP16 MFMA products, FP32 accumulation; also
    a field whose values span 12 decades (the per-chunk scaling keeps the small
    columns' relative accuracy)."""
    say(f"f16 split-2 modeprod n={n} rotate={rotate}")
    rng = np.random.default_rng(n + 17)
    Vm, Vim, D = O.eigbasis(n, 3)
    g = pyetd.Etd.from_dict(dict(small_params(n=n, nac=1), precision=2))
    for A in (Vim, Vm):
        u = rng.standard_normal(n ** 3)
        got = g.modeprod(A, u, rotate)
        ref = O.modeprod(A, u, rotate)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        say(f"  rel err {err:.2e}")
        assert err < 2e-6
        if n <= 256:
            # the kernel the step uses at n <= 256: contraction over the slowest
            # axis, i.e. the oracle's fastest-axis product of the field with its
            # slowest axis moved to the front
            got = g.modeprod(A, u, rotate, slowest=True)
            ut = np.ascontiguousarray(u.reshape(n, n, n).transpose(1, 2, 0)).reshape(-1)
            ref = O.modeprod(A, ut, not rotate)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            say(f"  slowest-axis rel err {err:.2e}")
            assert err < 2e-6
    # columns (lines along x) scaled over 12 decades: each column's own error
    # stays relative to that column
    scale = 10.0 ** rng.uniform(-12, 0, size=(n, n, 1))
    u = (rng.standard_normal((n, n, n)) * scale).ravel()
    got = g.modeprod(Vim, u, False)
    ref = O.modeprod(Vim, u, False)
    colmax = np.abs(ref.reshape(n, n, n)).max(axis=2, keepdims=True)
    err = (np.abs(got - ref).reshape(n, n, n) / colmax).max()
    say(f"  per-column rel err over 12 decades {err:.2e}")
    assert err < 2e-6
    g.close()


@pytest.mark.parametrize("case", MIXED_CASES, ids=lambda c: "n{n}_ac{nac}_ch{nch}_etd2{etd2}_el{elastic}_an{aniso}".format(**c))
def test_f16_mixed_step_matches_oracle(case):
    say(f"f16 mixed step {case}")
    c = dict(case)
    steps = c.pop("steps")
    prm = small_params(dt=0.02, **c)
    o, g = pair(prm, noise_fields(prm, amp=0.3), precision=2)
    for _ in range(steps):
        o.step()
        g.step()
    worst = max(np.abs(g.get_field(v) - o.get_field(v)).max() for v in range(prm["nac"] + prm["nch"]))
    say(f"  max |d| = {worst:.3e}")
    assert worst < 1e-5


@pytest.mark.parametrize("kernel", ["0", "5"])
@pytest.mark.parametrize("n,nac,nch", [(20, 2, 1), (36, 2, 1), (100, 4, 1), (132, 6, 1)])
def test_f16_three_slot_pipeline_equals_default(n, nac, nch, kernel, monkeypatch):
    """The FP16 split-2 kernels whose field chunks are prefetched two chunks
    ahead (the default, with whole-line field copies into an XOR-swizzled LDS
    layout; ETD_SPLIT_KERNEL=5, 32-byte copies) compute exactly the products
    of the two-slot kernel (ETD_SPLIT_KERNEL=7): bit-identical fields after
    three elastic steps.  The cases cover one K
    chunk per tile (n = 20), the partial last column block, and workgroups that
    run several tiles in a row (n = 132: 14 fields x 69 column blocks over 256
    workgroups), where the copy pipeline crosses tile boundaries."""
    say(f"f16 three-slot pipeline kernel={kernel} n={n} nac={nac} nch={nch}")
    prm = small_params(n=n, nac=nac, nch=nch, elastic=1, dt=0.02)
    fields = noise_fields(prm, amp=0.3)
    Vm, Vim, D = O.eigbasis(n, 3)

    def run(kernel):
        if kernel is None:
            monkeypatch.delenv("ETD_SPLIT_KERNEL", raising=False)
        else:
            monkeypatch.setenv("ETD_SPLIT_KERNEL", kernel)
        g = pyetd.Etd(pyetd.params_from_dict(dict(prm, precision=2)), Vm, Vim, D)
        for v, f in enumerate(fields):
            g.set_field(v, f)
        g.step(3)
        out = [g.get_field(v) for v in range(nac + nch)]
        g.close()
        return out

    a = run("7")
    b = run(kernel)
    for v, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x, y), f"field {v}: max |d| {np.abs(x - y).max():.3e}"
    monkeypatch.delenv("ETD_SPLIT_KERNEL", raising=False)


@pytest.mark.parametrize("n,nac,nch", [(20, 2, 1), (36, 2, 1), (132, 6, 1), (260, 2, 1)])
def test_bf16_whole_line_copies_equal_32byte_copies(n, nac, nch, monkeypatch):
    """The BF16 split-3 kernel's default field copies (whole 128-byte column
    runs into an XOR-swizzled LDS layout) give exactly the products of its
    32-byte copies into [k/8][col][8 k] (ETD_SPLIT_KERNEL=9): bit-identical
    fields after three elastic steps, with 8 waves (n <= 224) and with the
    4-wave tiling of larger n (260)."""
    say(f"bf16 whole-line copies n={n}")
    prm = small_params(n=n, nac=nac, nch=nch, elastic=1, dt=0.02)
    fields = noise_fields(prm, amp=0.3)
    Vm, Vim, D = O.eigbasis(n, 3)

    def run(kernel):
        monkeypatch.setenv("ETD_SPLIT_KERNEL", kernel)
        g = pyetd.Etd(pyetd.params_from_dict(dict(prm, precision=1)), Vm, Vim, D)
        for v, f in enumerate(fields):
            g.set_field(v, f)
        g.step(3)
        out = [g.get_field(v) for v in range(nac + nch)]
        g.close()
        return out

    a = run("9")
    b = run("0")
    for v, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x, y), f"field {v}: max |d| {np.abs(x - y).max():.3e}"
    monkeypatch.delenv("ETD_SPLIT_KERNEL", raising=False)

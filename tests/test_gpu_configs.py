"""GPU parity at the sizes of the benchmark configurations, through the C ABI.

The kernels are templated on the number of 16-row tiles, MT = ceil(n / 16),
so the instances that carry the 192^3 workload (n = 196, MT = 13) and the
256^3 one (n = 260, MT = 17) are only exercised at those sizes:

- mode products at n = 196 against the FP64 oracle, for the FP64 kernel,
  FP16 split-2 and BF16 split-3, by V and by V^-1, rotated or not;
- the 192^3 x (12 AC + 1 CH) beta -> alpha workload with cubic elasticity,
  two full ETD1 steps against the oracle: FP64 within 1e-10, both mixed
  precisions within 1e-5;
- the 256^3 x 4 coarsening workload (seeded nuclei), mixed FP16 split-2 as
  configured, two steps against the oracle within 1e-5;
- a 64-instance ensemble drawn as the ensemble configuration draws it
  (edges from {32, 48, 64, 80, 96}, a2 in [0.1, 0.4], eigenstrain scale in
  [0.01, 0.04], 4 variants, mixed, autotuned kernels), replayed as one
  hipGraph and checked bit for bit against stepping every instance alone,
  and one instance per edge against the FP64 oracle.

The oracle runs on the host cores (OpenMP), about 3 s per 192^3 step.
"""
import sys

import numpy as np
import pytest

import oracle as O
import pyetd
from pyetd import ensemble_c4 as C4
from pyetd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision", [0, 1, 2], ids=["fp64", "bf16x3", "fp16x2"])
@pytest.mark.parametrize("rotate", [True, False])
def test_modeprod_n196(precision, rotate):
    n = 196
    say(f"modeprod n={n} precision={precision} rotate={rotate}")
    rng = np.random.default_rng(196 + precision)
    Vm, Vim, D = O.eigbasis(n, 3)
    g = pyetd.Etd.from_dict(dict(W.scaled(W.beta_alpha_192(), n, nac=1, nch=0, elastic=0), precision=precision))
    for A in (Vim, Vm):
        u = rng.standard_normal(n ** 3)
        got = g.modeprod(A, u, rotate)
        ref = O.modeprod(A, u, rotate)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        say(f"  rel err {err:.2e}")
        assert err < (1e-13 if precision == 0 else 2e-6)
        if precision == 2:
            # the step's kernel at this size: contraction over the slowest axis
            # (the oracle's fastest-axis product of the transposed field)
            got = g.modeprod(A, u, rotate, slowest=True)
            ut = np.ascontiguousarray(u.reshape(n, n, n).transpose(1, 2, 0)).reshape(-1)
            ref = O.modeprod(A, ut, not rotate)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            say(f"  slowest-axis rel err {err:.2e}")
            assert err < 2e-6
    g.close()


@pytest.fixture(scope="module")
def alpha192():
    """The 192^3 x 13 workload: initial fields and the oracle after 2 steps."""
    prm = W.beta_alpha_192()
    f = W.initial_fields(prm)
    say(f"oracle: 2 steps of {prm['name']} on {O.threads()} host threads")
    o = O.Oracle(prm)
    for v, x in enumerate(f):
        o.set_field(v, x)
    o.step(2)
    ref = [o.get_field(v) for v in range(prm["nac"] + prm["nch"])]
    del o
    return prm, f, ref


@pytest.mark.parametrize("precision", [0, 1, 2], ids=["fp64", "bf16x3", "fp16x2"])
def test_beta_alpha_192_steps_match_oracle(alpha192, precision):
    prm, f, ref = alpha192
    say(f"192^3 x 13 elastic ETD1, 2 steps, precision={precision}")
    Vm, Vim, D = O.eigbasis(prm["n"], 3)
    g = pyetd.Etd(pyetd.params_from_dict(dict(prm, precision=precision)), Vm, Vim, D)
    for v, x in enumerate(f):
        g.set_field(v, x)
    g.step(2)
    nac = prm["nac"]
    deta = max(float(np.abs(g.get_field(v) - ref[v]).max()) for v in range(nac))
    dc = float(np.abs(g.get_field(nac) - ref[nac]).max())
    g.close()
    say(f"  max |d eta| = {deta:.3e}, max |d c| = {dc:.3e}")
    tol = 1e-10 if precision == 0 else 1e-5
    assert deta < tol and dc < tol


def test_beta_omega_256_coarsening_steps_match_oracle():
    prm = W.beta_omega_256_coarsening()  # mixed FP16 split-2, as configured
    say(f"256^3 x 4 nucleated, 2 steps, precision={prm['precision']}")
    f = W.initial_fields(prm)
    o = O.Oracle(dict(prm, precision=0))
    for v, x in enumerate(f):
        o.set_field(v, x)
    o.step(2)
    Vm, Vim, D = O.eigbasis(prm["n"], 3)
    g = pyetd.Etd(pyetd.params_from_dict(prm), Vm, Vim, D)
    for v, x in enumerate(f):
        g.set_field(v, x)
    g.step(2)
    worst = max(float(np.abs(g.get_field(v) - o.get_field(v)).max()) for v in range(4))
    g.close()
    say(f"  max |d eta| = {worst:.3e}")
    assert worst < 1e-5


def test_ensemble_64_instances_equals_solo_and_oracle():
    prms = C4.instances(64, 4000)
    sizes = sorted({p["n"] for p in prms})
    say(f"64-instance ensemble, extended edges {sizes}")
    assert len(prms) == 64 and set(sizes) <= {36, 52, 68, 84, 100}
    tunings = {n: pyetd.autotune(next(p for p in prms if p["n"] == n), steps=2) for n in sizes}
    members = [C4.make(p, tunings[p["n"]]) for p in prms]
    solo = [C4.make(p, tunings[p["n"]], carry=False) for p in prms]  # as the graph: u transformed each step
    K = 6  # a multiple of the felas ring period 3 (nac = 4, ETD1)
    E = pyetd.Ensemble(members, K)
    for rep in range(2):
        E.run(1)
        for s in solo:
            s.step(K)
        for i, (a, b) in enumerate(zip(members, solo)):
            for v in range(4):
                assert np.array_equal(a.get_field(v), b.get_field(v)), (rep, i, v)
        say(f"  replay {rep}: 64 instances bit-identical to solo stepping")
    E.close()
    for n in sizes:
        i = next(k for k, p in enumerate(prms) if p["n"] == n)
        o = O.Oracle(dict(prms[i], precision=0))
        for v, x in enumerate(W.initial_fields(prms[i])):
            o.set_field(v, x)
        o.step(2 * K)
        d = max(float(np.abs(members[i].get_field(v) - o.get_field(v)).max()) for v in range(4))
        say(f"  instance {i} (n = {n}) vs FP64 oracle after {2 * K} steps: {d:.2e}")
        assert d < 1e-5
    for g in members + solo:
        g.close()

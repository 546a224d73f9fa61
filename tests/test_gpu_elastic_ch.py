"""GPU parity of the elastic force on the CH fields (ELASTIC = 2,
elastic.c:726-760 with calc.c:90-101 and the outside chemical potential of
ScLETD.h:153-156) and of ELASTIC = 3 (no force), through the C ABI, against
the oracle (pinned in tests/test_oracle_elastic_ch.py)."""
import sys

import numpy as np
import pytest

import oracle as O
import pyetd
from conftest import small_params

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def ch_params(elastic=2, nac=2, nch=3, n=20, **kw):
    prm = small_params(n=n, nac=max(nac, nch), elastic=1, seed=31, dt=0.002, **kw)
    prm.update(nac=nac, nch=nch, elastic=elastic)
    return prm


def fields(prm, seed=32):
    rng = np.random.default_rng(seed)
    n = prm["n"]
    return ([rng.uniform(-0.3, 0.3, size=(n, n, n)) for _ in range(prm["nac"])]
            + [0.2 + 0.1 * rng.standard_normal((n, n, n)) for _ in range(prm["nch"])])


def pair(prm, f, precision=0):
    o = O.Oracle(prm)
    Vm, Vim, D = O.eigbasis(prm["n"], 3 if prm.get("periodic", 1) else 0)
    g = pyetd.Etd(pyetd.params_from_dict(dict(prm, precision=precision)), Vm, Vim, D)
    for v, x in enumerate(f):
        o.set_field(v, x)
        g.set_field(v, x)
    return o, g


def test_ch_force_matches_oracle_over_two_evaluations():
    say("CH elastic force")
    prm = ch_params()
    o, g = pair(prm, fields(prm))
    for ev in range(2):
        o.transfer()
        g.transfer()
        o.elastic()
        got = g.elastic()
        for q in range(3):
            ref = o.felas(q)
            assert np.abs(got[q] - ref).max() < 1e-12 * max(1e-3, np.abs(ref).max()), (ev, q)
        for v in range(5):  # change the fields for the second evaluation
            x = o.get_field(v) * 1.1 + 0.01
            o.set_field(v, x)
            g.set_field(v, x)
    g.close()


@pytest.mark.parametrize("varmob", [0, 1])
def test_ch_nonlinear_with_force_matches_oracle(varmob):
    say(f"CH nonlinear with force, varmob={varmob}")
    prm = ch_params(nac=1 if varmob else 2)
    prm["varmob"] = varmob
    f = fields(prm)
    if varmob:
        f[0] = np.abs(f[0]) * 2
    o, g = pair(prm, f)
    o.transfer()
    g.transfer()
    o.elastic()
    g.elastic(want=False)
    for v in range(prm["nac"], prm["nac"] + 3):
        ref = o.nonlinear(v)
        got = g.nonlinear(v)
        assert np.abs(got - ref).max() < 1e-11 * max(1.0, np.abs(ref).max()), v
    g.close()


CASES = [
    ("etd1", dict()),
    ("etd2", dict(etd2=1)),
    ("physical_etd2", dict(etd2=1, periodic=0)),
]


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("precision", [0, 2])
def test_steps_with_ch_force_match_oracle(name, kw, precision):
    say(f"CH force steps {name} precision={precision}")
    kw = dict(kw)
    periodic = kw.pop("periodic", 1)
    prm = ch_params(**kw)
    prm["periodic"] = periodic
    f = fields(prm)
    if precision:
        f = [x.astype(np.float32).astype(np.float64) for x in f]
    o, g = pair(prm, f, precision)
    for s in range(3):
        o.step()
        g.step()
    worst = max(float(np.abs(g.get_field(v) - o.get_field(v)).max()) for v in range(5))
    g.close()
    say(f"  max |d| = {worst:.3e}")
    assert worst < (1e-10 if precision == 0 else 1e-5)


def test_elastic_both_is_force_free():
    say("ELASTIC = 3")
    out = {}
    for el in (0, 3):
        prm = ch_params(elastic=el)
        f = fields(prm)
        Vm, Vim, D = O.eigbasis(prm["n"], 3)
        g = pyetd.Etd(pyetd.params_from_dict(prm), Vm, Vim, D)
        for v, x in enumerate(f):
            g.set_field(v, x)
        g.step(3)
        out[el] = [g.get_field(v) for v in range(5)]
        g.close()
    for v in range(5):
        assert np.array_equal(out[0][v], out[3][v])

"""The AC spectrum carried between fixed steps (csrc/etd_carry.hip): every step
after the first transforms only [u_CH | N(u)] and obtains the AC spectrum of
the ghost-filled fields as the carried spectrum plus the transform of the
changes the ghost fill made (transfer, trans.c:1939-1996).  In exact
arithmetic that is the forward transform of u itself, so:

- steps with the carried spectrum equal steps that transform u (ETD_CARRY=0)
  to roundoff, and both equal the oracle within the step tolerances (FP64
  1e-10; mixed 1e-5), for ETD1 / ETD2, elasticity, both anisotropic
  stencils, CH fields and physical boundaries;
- setting a field between steps drops the carried spectrum."""
import os
import sys

import numpy as np
import pytest

import oracle as O
import pyetd
from conftest import OMEGA_LANDAU, noise_fields, small_params

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu(prm, carry):
    if not carry:
        os.environ["ETD_CARRY"] = "0"
    try:
        Vm, Vim, D = O.eigbasis(prm["n"], 3 if prm.get("periodic", 1) else 0)
        return pyetd.Etd(pyetd.params_from_dict(prm), Vm, Vim, D)
    finally:
        os.environ.pop("ETD_CARRY", None)


CASES = [
    dict(n=16, nac=3, nch=1, etd2=0, elastic=1, aniso=0, precision=0, steps=6),
    dict(n=16, nac=3, nch=1, etd2=1, elastic=1, aniso=1, precision=0, steps=5),
    dict(n=20, nac=2, nch=0, etd2=0, elastic=0, aniso=2, precision=0, steps=5),
    dict(n=16, nac=3, nch=1, etd2=0, elastic=1, aniso=0, precision=2, steps=6),
    dict(n=20, nac=4, nch=1, etd2=1, elastic=1, aniso=1, precision=2, steps=5),
    dict(n=16, nac=2, nch=1, etd2=1, elastic=0, aniso=0, precision=0, steps=5, periodic=0),
    dict(n=20, nac=3, nch=1, etd2=0, elastic=1, aniso=0, precision=1, steps=5),  # BF16 split-3 products
    dict(n=16, nac=1, nch=0, etd2=0, elastic=0, aniso=0, precision=0, steps=6),  # one AC field
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "n{n}_ac{nac}_ch{nch}_etd2{etd2}_el{elastic}_an{aniso}_p{precision}".format(**c)
                         + ("_phys" if c.get("periodic", 1) == 0 else ""))
def test_carried_steps_match_transformed_steps_and_oracle(case):
    say(f"carry {case}")
    c = dict(case)
    steps = c.pop("steps")
    precision = c.pop("precision")
    periodic = c.pop("periodic", 1)
    prm = dict(small_params(dt=0.02, landau=OMEGA_LANDAU, KE=42.0, h=(1.0, 0.9, 1.1), **c), precision=precision,
               periodic=periodic)
    f = noise_fields(prm, amp=0.3)
    if precision:
        f = [x.astype(np.float32).astype(np.float64) for x in f]
    a, b = gpu(prm, True), gpu(prm, False)
    o = O.Oracle(dict(prm, precision=0))
    for v, x in enumerate(f):
        a.set_field(v, x)
        b.set_field(v, x)
        o.set_field(v, x)
    V = prm["nac"] + prm["nch"]
    for s in range(steps):
        a.step()
        b.step()
        o.step()
    d_ab = max(np.abs(a.get_field(v) - b.get_field(v)).max() for v in range(V))
    d_ao = max(np.abs(a.get_field(v) - o.get_field(v)).max() for v in range(V))
    say(f"  carried vs transformed {d_ab:.3e}, carried vs oracle {d_ao:.3e}")
    if precision == 0:
        assert d_ab < 1e-12 and d_ao < 1e-10
    else:
        assert d_ab < 1e-5 and d_ao < 1e-5
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [0, 2])
def test_carried_steps_on_an_edge_wider_than_a_block(precision):
    """n = 264 > 256: the correction covers every column i^, not only the
    first block's worth (a GPU-only comparison, the oracle is too slow here)."""
    say(f"carry n = 264, precision {precision}")
    prm = dict(small_params(n=264, nac=2, nch=0, etd2=0, elastic=0, dt=0.02, landau=OMEGA_LANDAU), precision=precision)
    f = noise_fields(prm, amp=0.3)
    a, b = gpu(prm, True), gpu(prm, False)
    for v, x in enumerate(f):
        a.set_field(v, x)
        b.set_field(v, x)
    a.step(3)
    b.step(3)
    d_ab = max(np.abs(a.get_field(v) - b.get_field(v)).max() for v in range(2))
    say(f"  carried vs transformed {d_ab:.3e}")
    assert d_ab < (1e-12 if precision == 0 else 1e-5)
    a.close()
    b.close()


def test_carried_steps_do_not_drift_over_a_long_run():
    """300 FP64 steps with elasticity and a CH field: each carried step
    starts from the spectrum the last backward transform consumed, so the
    carried and the transformed runs stay together at roundoff instead of
    drifting apart."""
    say("carry over 300 steps")
    prm = small_params(n=20, nac=3, nch=1, etd2=0, elastic=1, aniso=1, dt=0.02, landau=OMEGA_LANDAU, KE=42.0)
    f = noise_fields(prm, amp=0.3, seed=11)
    a, b = gpu(prm, True), gpu(prm, False)
    for v, x in enumerate(f):
        a.set_field(v, x)
        b.set_field(v, x)
    worst = []
    for rep in range(3):
        a.step(100)
        b.step(100)
        worst.append(max(float(np.abs(a.get_field(v) - b.get_field(v)).max()) for v in range(4)))
    say(f"  carried vs transformed after 100, 200, 300 steps: {worst}")
    assert max(worst) < 1e-11
    a.close()
    b.close()


def test_setting_a_field_drops_the_carried_spectrum():
    """Step, replace the fields, step again: the same as a fresh context that
    starts from the replaced fields."""
    say("carry dropped by set_field")
    prm = small_params(n=16, nac=3, nch=1, etd2=0, elastic=1, dt=0.02)
    f0 = noise_fields(prm, amp=0.3, seed=1)
    f1 = noise_fields(prm, amp=0.3, seed=2)
    a = gpu(prm, True)
    for v, x in enumerate(f0):
        a.set_field(v, x)
    a.step(2)
    for v, x in enumerate(f1):
        a.set_field(v, x)
    a.step(2)
    # the force ring's ghost lags carry the history of earlier evaluations,
    # so the reference is the oracle taken through the same sequence
    o = O.Oracle(prm)
    for v, x in enumerate(f0):
        o.set_field(v, x)
    o.step(2)
    for v, x in enumerate(f1):
        o.set_field(v, x)
    o.step(2)
    worst = max(np.abs(a.get_field(v) - o.get_field(v)).max() for v in range(4))
    say(f"  max |d| = {worst:.3e}")
    assert worst < 1e-10
    a.close()

"""GPU parity of the beta -> alpha physics and the variable-mobility CH,
through the C ABI, against the oracle (oracle/etd_oracle.c, pinned in
tests/test_oracle_alpha.py):

- the CALPHAD driving force (anisotropic_hip.c:2002-2188) with and without
  the grain mask, alone and with each gradient term;
- the compact Lambda_ij d_i d_j stencil (anisotropic_hip.c:528-647), any
  Lambda, non-unit spacing;
- the variable-mobility CH term (mobility.c:838-922);
- full ETD1 / ETD2 steps with these, FP64 within 1e-10, mixed within 1e-5."""
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


def pair(prm, fields, mask=None):
    o = O.Oracle(dict(prm, mask=mask) if mask is not None else prm)
    Vm, Vim, D = O.eigbasis(prm["n"], 3)
    g = pyetd.Etd(pyetd.params_from_dict(prm), Vm, Vim, D)
    if mask is not None:
        g.set_mask(mask)
    for v, f in enumerate(fields):
        o.set_field(v, f)
        g.set_field(v, f)
    return o, g


def alpha_params(n=16, nac=6, aniso=0, elastic=0, etd2=0, seed=1):
    prm = dict(small_params(n=n, nac=nac, aniso=1 if aniso else 0, elastic=elastic, etd2=etd2, seed=seed,
                            h=(0.9, 1.0, 1.1), kk=(1.0, 1.2, 0.8), dt=0.002),
               free_energy=1, LE=2.0, KE=3.0, epn2=0.048)
    prm["aniso"] = aniso
    return prm


def eta_fields(prm, seed=7):
    rng = np.random.default_rng(seed)
    n = prm["n"]
    return [rng.uniform(-0.1, 1.1, size=(n, n, n)) for _ in range(prm["nac"] + prm["nch"])]


@pytest.mark.parametrize("aniso", [0, 1, 2])
@pytest.mark.parametrize("masked", [False, True])
def test_calphad_nonlinear_matches_oracle(aniso, masked):
    say(f"calphad nonlinear aniso={aniso} mask={masked}")
    prm = alpha_params(aniso=aniso)
    f = eta_fields(prm)
    mask = (np.random.default_rng(2).uniform(size=(16, 16, 16)) > 0.3).astype(np.int32) if masked else None
    o, g = pair(prm, f, mask)
    o.transfer()
    g.transfer()
    for p in range(prm["nac"]):
        ref = o.nonlinear(p)
        got = g.nonlinear(p)
        assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()), p
    g.close()


@pytest.mark.parametrize("n", [12, 21])
def test_compact_stencil_nonlinear_matches_oracle(n):
    say(f"compact stencil n={n}")
    prm = dict(small_params(n=n, nac=3, aniso=1, h=(0.9, 1.0, 1.2), kk=(1.0, 1.3, 0.7)), aniso=2)
    o, g = pair(prm, noise_fields(prm, amp=0.5))
    o.transfer()
    g.transfer()
    for p in range(3):
        ref = o.nonlinear(p)
        got = g.nonlinear(p)
        assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()), p
    g.close()


def vm_params(n=14, etd2=0, precision=0):
    prm = dict(small_params(n=n, nac=2, nch=3, etd2=etd2, h=(0.9, 1.0, 1.1), dt=0.002), varmob=1,
               KCI=0.5, precision=precision)
    return prm


def vm_fields(prm, seed=9):
    rng = np.random.default_rng(seed)
    n = prm["n"]
    eta = [rng.uniform(0, 1, size=(n, n, n)) for _ in range(prm["nac"])]
    c = [0.3 + 0.05 * rng.standard_normal((n, n, n)) for _ in range(3)]
    return eta + c


def test_variable_mobility_nonlinear_matches_oracle():
    say("variable mobility nonlinear")
    prm = vm_params()
    o, g = pair(prm, vm_fields(prm))
    o.transfer()
    g.transfer()
    for v in range(2, 5):
        ref = o.nonlinear(v)
        got = g.nonlinear(v)
        assert np.abs(got - ref).max() < 1e-11 * max(1.0, np.abs(ref).max()), v
    g.close()


STEPS = [
    ("calphad_aniso2_mask_elastic_etd2", lambda: alpha_params(n=20, nac=4, aniso=2, elastic=1, etd2=1), True),
    ("calphad_aniso1", lambda: alpha_params(n=16, nac=6, aniso=1), False),
    ("calphad_etd1_mask", lambda: alpha_params(n=16, nac=6), True),
    ("varmob_etd2", lambda: vm_params(n=16, etd2=1), False),
    ("varmob_etd1", lambda: vm_params(n=16), False),
]


@pytest.mark.parametrize("name,make,masked", STEPS, ids=[s[0] for s in STEPS])
@pytest.mark.parametrize("precision", [0, 2])
def test_steps_match_oracle(name, make, masked, precision):
    say(f"steps {name} precision={precision}")
    prm = make()
    n = prm["n"]
    f = vm_fields(prm) if prm.get("varmob") else eta_fields(prm)
    if precision:
        f = [x.astype(np.float32).astype(np.float64) for x in f]
    mask = (np.random.default_rng(3).uniform(size=(n, n, n)) > 0.3).astype(np.int32) if masked else None
    o = O.Oracle(dict(prm, mask=mask) if mask is not None else prm)
    Vm, Vim, D = O.eigbasis(n, 3)
    g = pyetd.Etd(pyetd.params_from_dict(dict(prm, precision=precision)), Vm, Vim, D)
    if mask is not None:
        g.set_mask(mask)
    for v, x in enumerate(f):
        o.set_field(v, x)
        g.set_field(v, x)
    for s in range(3):
        o.step()
        g.step()
    V = prm["nac"] + prm["nch"]
    worst = max(float(np.abs(g.get_field(v) - o.get_field(v)).max()) for v in range(V))
    scale = max(float(np.abs(o.get_field(v)).max()) for v in range(V))
    g.close()
    say(f"  max |d| = {worst:.3e} (field scale {scale:.2f})")
    assert worst < (1e-10 if precision == 0 else 1e-5 * max(1.0, scale))


def test_adaptive_refuses_the_alpha_physics():
    prm = alpha_params(n=12, nac=2)
    g = pyetd.Etd.from_dict(prm)
    with pytest.raises(pyetd.EtdError) as e:
        g.run_adaptive(t_end=0.01, dt0=0.002, tol=1e-4)
    assert e.value.code == pyetd.ERR_UNSUPPORTED
    g.close()


def test_variable_mobility_needs_three_ch_fields():
    prm = dict(small_params(n=12, nac=1, nch=2), varmob=1)
    with pytest.raises(pyetd.EtdError):
        pyetd.Etd.from_dict(prm)

"""GPU parity for physical (non-periodic) boundaries on one box, through the
C ABI, against the oracle (pinned in tests/test_oracle_boundaries.py):
boundary id 0 eigenbasis, no ghost fill, no outside values, omega's
low-mirror / high-zero stencils, alpha's both-sides mirror, felas 0 outside
the interior, check_soln_new over [0, n - g); nonlinear terms within 1e-12,
three steps within 1e-10 (FP64) and 1e-5 (mixed), and the drop-in driver on
a periodic = 0 ini."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
import pyetd
from conftest import OMEGA_LANDAU, ROOT, noise_fields, small_params

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def phys_params(**kw):
    extra = {k: kw.pop(k) for k in ("free_energy", "varmob", "precision", "LE", "KE", "epn2") if k in kw}
    aniso = kw.pop("aniso", 0)
    prm = small_params(aniso=1 if aniso else 0, **kw)
    prm.update(extra)
    prm["aniso"] = aniso
    prm["periodic"] = 0
    return prm


def pair(prm, fields):
    n = prm["n"]
    o = O.Oracle(prm)
    Vm, Vim, D = O.eigbasis(n, 0)
    g = pyetd.Etd(pyetd.params_from_dict(prm), Vm, Vim, D)
    for v, f in enumerate(fields):
        o.set_field(v, f)
        g.set_field(v, f)
    return o, g


@pytest.mark.parametrize("aniso", [0, 1, 2])
def test_ac_and_ch_nonlinear_match_oracle(aniso):
    say(f"physical nonlinear aniso={aniso}")
    prm = phys_params(n=14, nac=3, nch=1, aniso=aniso, h=(0.9, 1.0, 1.2), kk=(1.0, 1.3, 0.7))
    o, g = pair(prm, noise_fields(prm, amp=0.5, ch_amp=0.2))
    o.transfer()
    g.transfer()
    for v in range(4):
        ref = o.nonlinear(v)
        got = g.nonlinear(v)
        assert np.abs(got - ref).max() < 1e-11 * max(1.0, np.abs(ref).max()), v
    g.close()


def test_variable_mobility_nonlinear_matches_oracle():
    say("physical variable mobility")
    prm = phys_params(n=14, nac=1, nch=3, varmob=1, h=(0.9, 1.0, 1.1))
    rng = np.random.default_rng(4)
    f = [rng.uniform(0, 1, size=(14, 14, 14))] + [0.3 + 0.05 * rng.standard_normal((14, 14, 14)) for _ in range(3)]
    o, g = pair(prm, f)
    for v in range(1, 4):
        ref = o.nonlinear(v)
        got = g.nonlinear(v)
        assert np.abs(got - ref).max() < 1e-11 * max(1.0, np.abs(ref).max()), v
    g.close()


def test_elastic_force_is_zero_outside_the_interior():
    say("physical elastic")
    prm = phys_params(n=20, nac=3, elastic=1)
    o, g = pair(prm, noise_fields(prm, amp=0.4))
    o.elastic()
    got = g.elastic()
    for p in range(3):
        ref = o.felas(p)
        assert np.abs(got[p] - ref).max() < 1e-12 * max(1e-3, np.abs(ref).max()), p
    g.close()


@pytest.mark.parametrize("n", [14, 68])
def test_diag_matches_oracle(n):
    say(f"physical diag n={n}")
    prm = phys_params(n=n, nac=3, nch=1, aniso=1, landau=OMEGA_LANDAU, h=(0.9, 1.0, 1.2))
    o, g = pair(prm, noise_fields(prm, amp=0.5))
    st_o, cp_o = o.diag()
    st_g, cp_g = g.diag()
    g.close()
    assert np.array_equal(st_o[:, :2], st_g[:, :2])
    assert np.abs(st_o[:, 2:] - st_g[:, 2:]).max() < 1e-12 * np.abs(st_o[:, 2:]).max()
    assert abs(cp_o - cp_g) < 1e-12 * abs(cp_o)


CASES = [
    ("ac_ch_aniso_elastic_etd2", dict(n=20, nac=3, nch=1, aniso=1, elastic=1, etd2=1)),
    ("ac_compact_etd1", dict(n=16, nac=2, aniso=2)),
    ("calphad_compact", dict(n=16, nac=4, aniso=2, free_energy=1, LE=2.0, KE=3.0, epn2=0.048)),
    ("varmob_etd2", dict(n=16, nac=1, nch=3, varmob=1, etd2=1)),
]


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("precision", [0, 2])
def test_steps_match_oracle(name, kw, precision):
    say(f"physical steps {name} precision={precision}")
    prm = phys_params(dt=0.002, **kw)
    n = prm["n"]
    rng = np.random.default_rng(6)
    if prm.get("varmob"):
        f = [rng.uniform(0, 1, size=(n, n, n))] + [0.3 + 0.05 * rng.standard_normal((n, n, n)) for _ in range(3)]
    else:
        f = noise_fields(prm, amp=0.3, ch_amp=0.1)
    if precision:
        f = [x.astype(np.float32).astype(np.float64) for x in f]
    o = O.Oracle(prm)
    Vm, Vim, D = O.eigbasis(n, 0)
    g = pyetd.Etd(pyetd.params_from_dict(dict(prm, precision=precision)), Vm, Vim, D)
    for v, x in enumerate(f):
        o.set_field(v, x)
        g.set_field(v, x)
    for s in range(3):
        o.step()
        g.step()
    V = prm["nac"] + prm["nch"]
    worst = max(float(np.abs(g.get_field(v) - o.get_field(v)).max()) for v in range(V))
    g.close()
    say(f"  max |d| = {worst:.3e}")
    assert worst < (1e-10 if precision == 0 else 1e-5)


def test_dropin_driver_on_a_physical_box(tmp_path):
    """etd_run on a periodic = 0 ini (no .dat files: the id 0 basis is generated,
    h = L / nx): the phi_ac CSV rows of 4 iterations against the oracle."""
    n, dt = 16, 2.0 ** -9
    d = str(tmp_path)
    ini = (f" {n}, {n}, {n}\n 1, 1, 1\n 2\n ./output/\n ./data/\n {dt!r}, {4 * dt!r}\n 1.0\n 0\n 2, 0\n 0, 0, ./r/\n"
           f" 0, 0, 0\n {float(n)!r}, {float(n)!r}, {float(n)!r}\n 0\n 0\n 0\n 1.0, 1.0, 1.0\n 0\n")
    os.makedirs(os.path.join(d, "output"))
    os.makedirs(os.path.join(d, "data"))
    open(os.path.join(d, "input5.ini"), "w").write(ini)
    open(os.path.join(d, "landau_input.txt"), "w").write("1.0 3.2 0.2 3.0 3.0\n")
    env = dict(os.environ, ETD_SEED="3")
    r = subprocess.run([os.path.join(ROOT, "energy-stability_amd", "etd_run"), "5", "."], cwd=d, env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [l.split(",") for l in open(os.path.join(d, "output", "phi_ac_0.csv")).read().split("\n")[1:] if l]
    assert len(rows) == 4
    # the same state: etd_init_noise seeds as the driver seeds them; h = L / nx = 1
    prm = dict(small_params(n=n, nac=2, dt=dt, KE=3.2, landau=(0.2, 3.0, 3.0)), periodic=0)
    g = pyetd.Etd.from_dict(prm)
    f0 = []
    for v in range(2):
        g.init_noise(v, 3 + 7919 * v, 0.0, 0.01)
        f0.append(g.get_field(v))
    g.close()
    o = O.Oracle(prm)
    for v in range(2):
        o.set_field(v, f0[v])
    for it, row in enumerate(rows):
        st, _ = o.diag()
        assert int(row[0]) == it
        assert row[1] == f"{st[0][0]:+1.15f}" and row[2] == f"{st[0][1]:+1.15f}"
        assert abs(float(row[3]) - st[0][2]) <= 1e-12 * abs(st[0][2]) + 1e-15
        assert abs(float(row[4]) - st[0][3]) <= 1e-12 * abs(st[0][3]) + 1e-15
        o.step()

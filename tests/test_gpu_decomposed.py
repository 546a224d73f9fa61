"""GPU parity of decomposed runs (omega's procs px,py,pz with periodic = 1:
every box on this GPU, exchanged as the reference's ranks exchange them,
trans.c:7-251,1939-1996, elastic.c:617-762), through the C ABI, against the
oracle's decomposed runs (orc_group, pinned in tests/test_oracle_decomposed.py):

- transfer bit for bit, the nonlinear terms within 1e-12 and the elastic
  force (every extended cell of every box) within 1e-12;
- steps (ETD1 / ETD2, elastic, both anisotropic stencils, CH) within 1e-10
  and check_soln_new summed over the boxes within 1e-12;
- the drop-in driver on the reference's own input64.ini (2 x 2 x 2 boxes of
  68^3) from the reference's sphere start: its first CSV row is the first
  line of the reference's shipped energy trace, and the next rows follow the
  oracle."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
import pyetd
from conftest import GOLDEN, OMEGA_LANDAU, small_params

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pair(prm, procs, seed=3, amp=0.3):
    """Oracle group and GPU context with the same boxes and fields (ghost
    cells random too: the exchange must overwrite or keep them as the
    reference does)."""
    P = procs[0] * procs[1] * procs[2]
    G = O.OracleGroup(prm, procs)
    Vm, Vim, D = O.eigbasis(prm["n"], 3)
    g = pyetd.Etd(pyetd.params_from_dict(dict(prm, procs=procs)), Vm, Vim, D)
    rng = np.random.default_rng(seed)
    n = prm["n"]
    for v in range(prm["nac"] + prm["nch"]):
        f = rng.uniform(-amp, amp, size=(P, n, n, n))
        if v >= prm["nac"]:
            f = 0.1 + 0.05 * f
        for b in range(P):
            G.box(b).set_field(v, f[b])
        g.set_field(v, f)
    return G, g


def test_transfer_and_nonlinear_match_oracle():
    procs = (2, 3, 2)
    say(f"decomposed transfer / nonlinear {procs}")
    prm = small_params(n=10, nac=2, nch=1, aniso=1, h=(0.9, 1.0, 1.1), landau=OMEGA_LANDAU, KE=42.0)
    G, g = pair(prm, procs)
    G.transfer()
    g.transfer()
    for v in range(3):
        got = g.get_field(v)
        for b in range(12):
            assert np.array_equal(got[b], G.box(b).get_field(v)), (v, b)
    for v in range(3):
        got = g.nonlinear(v)
        for b in range(12):
            ref = G.box(b).nonlinear(v)
            assert np.abs(got[b] - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()), (v, b)
    g.close()


def test_elastic_matches_oracle_over_two_evaluations():
    procs = (2, 2, 2)
    say(f"decomposed elastic {procs}")
    prm = small_params(n=12, nac=3, elastic=1)
    G, g = pair(prm, procs, amp=0.4)
    for ev in range(2):
        G.transfer()
        g.transfer()
        G.elastic()
        got = g.elastic()
        for p in range(3):
            for b in range(8):
                ref = G.box(b).felas(p)
                assert np.abs(got[p, b] - ref).max() < 1e-12 * max(1e-3, np.abs(ref).max()), (ev, p, b)
    g.close()


CASES = [
    dict(procs=(2, 1, 1), n=12, nac=2, nch=0, etd2=0, elastic=0, aniso=0, steps=4),
    dict(procs=(2, 2, 2), n=12, nac=3, nch=1, etd2=1, elastic=1, aniso=1, steps=3),
    dict(procs=(1, 2, 3), n=10, nac=2, nch=1, etd2=1, elastic=0, aniso=2, steps=3),
    dict(procs=(2, 2, 2), n=12, nac=2, nch=2, etd2=0, elastic=2, aniso=0, steps=3),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "p{}{}{}_n{n}_ac{nac}_ch{nch}_etd2{etd2}_el{elastic}_an{aniso}".format(
    *c["procs"], **c))
def test_steps_match_oracle(case):
    say(f"decomposed steps {case}")
    c = dict(case)
    procs, steps = c.pop("procs"), c.pop("steps")
    prm = small_params(dt=0.01, landau=OMEGA_LANDAU, KE=42.0, h=(1.0, 0.9, 1.1), **c)
    if prm["elastic"] == 2:  # eigenstrains per CH field
        prm["eps0"] = np.random.default_rng(4).uniform(-0.03, 0.03, size=(prm["nch"], 3, 3))
    G, g = pair(prm, procs, amp=0.2)
    V = prm["nac"] + prm["nch"]
    P = procs[0] * procs[1] * procs[2]
    for s in range(steps):
        G.step()
        g.step()
    worst = 0.0
    for v in range(V):
        got = g.get_field(v)
        for b in range(P):
            worst = max(worst, np.abs(got[b] - G.box(b).get_field(v)).max())
    say(f"  max |d| = {worst:.3e}")
    assert worst < 1e-10
    st_o, cp_o = G.diag()
    st_g, cp_g = g.diag()
    # the stepped fields agree to roundoff, so max and min do too (not bit for bit)
    assert np.abs(st_o[:, :2] - st_g[:, :2]).max() < 1e-10
    assert np.abs(st_o[:, 2:] - st_g[:, 2:]).max() < 1e-12 * np.abs(st_o[:, 2:]).max()
    assert abs(cp_o - cp_g) <= 1e-12 * abs(cp_o)
    g.close()


def test_decomposed_context_refuses_single_box_entry_points():
    prm = small_params(n=10, nac=1)
    g = pyetd.Etd(pyetd.params_from_dict(dict(prm, procs=(2, 1, 1))))
    with pytest.raises(pyetd.EtdError):
        g.run_adaptive(t_end=0.01, dt0=0.001, tol=1e-4)
    with pytest.raises(pyetd.EtdError):
        g.write_checkpoint(0, 0)
    g.close()


def test_dropin_driver_on_the_reference_input64(tmp_path):
    """etd_run 64 on the reference's omega inputs (procs 2,2,2, ETD2,
    elastic, anisotropic) from ETD_INIT=sphere, for three iterations."""
    say("etd_run 64, 2 x 2 x 2 boxes")
    d = str(tmp_path)
    for f in ("elastic_input.txt", "anisotropic_input.txt"):
        shutil.copy(os.path.join(GOLDEN, "omega", f), d)
    ini = open(os.path.join(GOLDEN, "omega", "input64.ini")).read()
    assert " 0.000001, 0.0066\n" in ini
    open(os.path.join(d, "input64.ini"), "w").write(ini.replace(" 0.000001, 0.0066\n", " 0.000001, 0.0000025\n"))
    os.makedirs(os.path.join(d, "output"))
    os.makedirs(os.path.join(d, "data"))
    env = dict(os.environ, ETD_INIT="sphere", ETD_OUTPUT_EVERY="0")
    r = subprocess.run([os.path.join(ROOT, "energy-stability_amd", "etd_run"), "64", "."], cwd=d, env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = [l.split(",") for l in open(os.path.join(d, "output", "phi_ac_3.csv")).read().split("\n")[1:] if l]
    assert [int(x[0]) for x in rows] == [0, 1, 2]
    ref0 = float(open(os.path.join(GOLDEN, "omega", "phi_ac_3_energy_head.txt")).readline().split()[1])
    assert abs(float(rows[0][4]) - ref0) <= 1e-13 * abs(ref0), (rows[0][4], ref0)
    # the oracle's decomposed run of the same inputs
    p = pyetd.parse_inputs(os.path.join(GOLDEN, "omega"), "64")
    prm = dict(n=68, g=2, nac=4, nch=0, etd2=1, elastic=1, aniso=1, dt=p.dt, epn2=p.epn2, h=(p.hx, p.hy, p.hz),
               kk=(p.kkx, p.kky, p.kkz), LE=p.LE, KE=p.KE, landau=(p.a2, p.a3, p.a4), elastic_scale=p.elastic_scale,
               C=np.array(p.C[:36]).reshape(6, 6), eps0=np.array(p.eps0[:36]).reshape(4, 3, 3),
               lam=np.array(p.lam[:36]).reshape(4, 3, 3))
    G = O.OracleGroup(prm, (2, 2, 2))
    n = 68
    for b in range(8):
        cx, cy, cz = b % 2, (b // 2) % 2, b // 4
        i = np.arange(n)
        r2 = ((cx * n + i - n)[None, None, :] ** 2 + (cy * n + i - n)[None, :, None] ** 2
              + (cz * n + i - n)[:, None, None] ** 2)
        for v in range(3):
            G.box(b).set_field(v, np.zeros((n, n, n)))
        G.box(b).set_field(3, np.where(r2 < 64, 1.0, 0.0))
    for it, row in enumerate(rows):
        st, _ = G.diag()
        say(f"  iter {it}: eng {row[4]} oracle {st[3][3]!r}")
        for col in (1, 2, 3, 4):
            assert abs(float(row[col]) - st[3][col - 1]) <= 1e-10 * max(1.0, abs(st[3][col - 1])), (it, col)
        G.step()

"""GPU tests of adaptive ETD and the learned dt controller, through the C ABI.

- The adaptive stepper with every attempt accepted and no extrapolation is plain
  stepping at dt/2, bit for bit.
- Its accept / reject decisions, error estimates, energies and states match
  the CPU restatement on the FP64 oracle (oracle/adaptive.py).
- The network's forward pass and gradient match the numpy oracle
  (oracle/mlp_oracle.py, itself pinned by finite differences); training is
  deterministic and reduces the loss.
- A network trained on plain adaptive runs keeps the guarantees (every
  accepted step within tolerance, monotone energy, t_end reached) and its
  gain over the plain controller is reported.
"""
import sys

import numpy as np
import pytest

import adaptive as AD
import mlp_oracle as M
import oracle as O
import pyetd
from conftest import noise_fields, small_params
from pyetd import controller as CT

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu(prm, fields, carry=True):
    Vm, Vim, D = O.eigbasis(prm["n"], 3)
    g = pyetd.Etd(pyetd.params_from_dict(prm), Vm, Vim, D, carry=carry)
    for v, f in enumerate(fields):
        g.set_field(v, f)
    return g


@pytest.mark.parametrize("precision,etd2,elastic", [(0, 0, 1), (2, 0, 1), (0, 1, 0), (0, 1, 1)])
def test_all_accepted_without_extrapolation_is_half_steps(precision, etd2, elastic):
    """With every attempt accepted and no extrapolation the adaptive stepper is
    plain stepping at dt / 2: bit for bit for ETD1, in FP64 and in the mixed
    path (whose full and first half step share one phi pass and whose CH mean
    restoration runs per state).  For ETD2 the attempt's full step makes two
    extra elastic evaluations, and the ghost cells of felas read the ring's
    older slots (the reference's shared transfer buffer), so the half steps
    see a different history in those cells: without elasticity ETD2 is bit
    for bit too, with it the difference stays at the ghost-cell level."""
    say(f"adaptive == half steps precision={precision} etd2={etd2} elastic={elastic}")
    dt = 2.0 ** -6
    prm = small_params(n=16, nac=3, nch=1, elastic=elastic, aniso=1, dt=dt, etd2=etd2)
    prm["precision"] = precision
    f = noise_fields(prm, amp=0.3)
    a = gpu(prm, f, carry=False)  # every attempt transforms u, as plain stepping without the carry does
    b = gpu(dict(prm, dt=dt / 2), f, carry=False)
    st, log = a.run_adaptive(log_cap=16, t_end=3 * dt, dt0=dt, tol=1e9, grow=1.0, shrink=1.0, extrapolate=0,
                             energy_test=0)
    assert st["accepted"] == 3 and st["rejected_err"] == 0 and st["step_evals"] == 9
    assert a.time == 3 * dt and len(log) == 3
    b.step(6)
    for v in range(4):
        if etd2 and elastic:
            d = np.abs(a.get_field(v) - b.get_field(v)).max()
            say(f"  field {v}: max |d| = {d:.2e}")
            assert d < 1e-6
        else:
            assert np.array_equal(a.get_field(v), b.get_field(v))


@pytest.mark.parametrize("etd2", [0, 1])
def test_one_extrapolated_attempt_is_two_half_steps_against_one_full_step(etd2):
    """One accepted attempt with local extrapolation is 2 u(dt/2, twice) -
    u(dt, once), bit for bit against plain stepping of both (FP64, no
    elasticity, every step transforming u): the full step and the first
    half step share their forward stage (and for ETD2 the predictor's), and
    each comes out as a step of its own would."""
    say(f"one extrapolated attempt etd2={etd2}")
    dt = 2.0 ** -6
    prm = small_params(n=16, nac=3, nch=1, elastic=0, aniso=1, dt=dt, etd2=etd2)
    f = noise_fields(prm, amp=0.3)
    a = gpu(prm, f, carry=False)
    half = gpu(dict(prm, dt=dt / 2), f, carry=False)
    full = gpu(prm, f, carry=False)
    st, _ = a.run_adaptive(log_cap=4, t_end=dt, dt0=dt, tol=1e9, grow=1.0, shrink=1.0, extrapolate=1,
                           energy_test=0)
    assert st["accepted"] == 1 and st["step_evals"] == 3
    half.step(2)
    full.step(1)
    for v in range(4):
        want = 2.0 * half.get_field(v) - full.get_field(v)
        assert np.array_equal(a.get_field(v), want), v
    for g in (a, half, full):
        g.close()


@pytest.mark.parametrize("elastic", [0, 1])
def test_plain_controller_matches_oracle_restatement(elastic):
    say(f"adaptive vs oracle restatement elastic={elastic}")
    prm = small_params(n=16, nac=3, nch=1, elastic=elastic, dt=0.01, KE=3.2, landau=(0.2, 3.0, 3.0))
    f = noise_fields(prm, amp=0.05, ch_amp=0.05)
    o = O.Oracle(prm)
    for v, x in enumerate(f):
        o.set_field(v, x)
    kw = dict(t_end=0.6, dt0=0.2, dt_min=1e-5, dt_max=0.3, tol=2e-4)
    t_o, rows_o, cnt_o = AD.run(o, **kw)
    g = gpu(prm, f)
    # the restatement tests the check_soln_new energy (no elastic term)
    st, log = g.run_adaptive(log_cap=1000, energy_elastic=0, energy_kind=0, **kw)
    say(f"  oracle {cnt_o}  gpu {st}")
    assert cnt_o["rejected_err"] > 0 and cnt_o["accepted"] > 3  # both branches exercised
    assert st["accepted"] == cnt_o["accepted"]
    assert st["rejected_err"] == cnt_o["rejected_err"] and st["rejected_energy"] == cnt_o["rejected_energy"]
    ro = np.array(rows_o)
    assert np.abs(log[:, 0] - ro[:, 0]).max() < 1e-12
    assert np.abs(log[:, 1] - ro[:, 1]).max() < 1e-12
    assert np.abs(log[:, 2] - ro[:, 2]).max() < 1e-9 * ro[:, 2].max()
    assert np.abs(log[:, 3] - ro[:, 3]).max() < 1e-10 * np.abs(ro[:, 3]).max()
    for v in range(4):
        assert np.abs(g.get_field(v) - o.get_field(v)).max() < 1e-10


@pytest.mark.parametrize("precision,extrapolate,etd2", [(0, 1, 0), (2, 1, 0), (0, 0, 0), (0, 1, 1), (2, 1, 1)])
def test_carried_attempts_match_transformed_attempts(precision, extrapolate, etd2):
    """Attempts that carry the AC spectrum (the second half step from the
    first's spectrum, the next attempt from spec or 2 spec - SP) make the same
    decisions as attempts that transform u, with fields equal to roundoff;
    rejections (which drop the carry) included; ETD1 and ETD2."""
    say(f"adaptive carried vs transformed precision={precision} extrapolate={extrapolate} etd2={etd2}")
    prm = dict(small_params(n=16, nac=3, nch=1, elastic=1, aniso=1, dt=0.01, KE=3.2, landau=(0.2, 3.0, 3.0),
                            etd2=etd2), precision=precision)
    f = noise_fields(prm, amp=0.05, ch_amp=0.05)
    if precision:
        f = [x.astype(np.float32).astype(np.float64) for x in f]
    kw = dict(t_end=0.6, dt0=0.2, dt_min=1e-5, dt_max=0.3, tol=2e-4, extrapolate=extrapolate)
    a, b = gpu(prm, f, carry=True), gpu(prm, f, carry=False)
    sa, la = a.run_adaptive(log_cap=1000, **kw)
    sb, lb = b.run_adaptive(log_cap=1000, **kw)
    say(f"  carried {sa}  transformed {sb}")
    assert sa["rejected_err"] + sa["rejected_energy"] > 0 and sa["accepted"] > 3
    for k in ("accepted", "rejected_err", "rejected_energy", "step_evals"):
        assert sa[k] == sb[k], k
    # dt follows sqrt(tol / err), and err is a difference of two states of
    # size tol: a roundoff-level change of the fields moves dt at 1e-11
    assert np.abs(la[:, 1] - lb[:, 1]).max() <= (1e-9 if precision == 0 else 1e-3) * lb[:, 1].max()
    d = max(np.abs(a.get_field(v) - b.get_field(v)).max() for v in range(4))
    say(f"  max |d| = {d:.2e}")
    assert d < (1e-11 if precision == 0 else 1e-5)
    # a fixed step after the run starts afresh (the run may end on 2 spec - SP)
    a.step(1)
    b.step(1)
    d = max(np.abs(a.get_field(v) - b.get_field(v)).max() for v in range(4))
    assert d < (1e-11 if precision == 0 else 1e-5)
    a.close()
    b.close()


def scheme_energy(prm, fields):
    """numpy restatement of the scheme's discrete free energy (interior,
    periodic forward differences; isotropic AC and CH)."""
    g = prm["g"]
    hx, hy, hz = prm["h"]
    kk = prm["kk"]
    a2, a3, a4 = prm["landau"]
    ep = prm["epn2"]
    inner = [f[g:-g, g:-g, g:-g] for f in fields]  # [z][y][x]
    e = 0.0
    s2 = np.zeros_like(inner[0])
    s3 = np.zeros_like(inner[0])
    for p, f in enumerate(inner):
        dx = (np.roll(f, -1, 2) - f) / hx
        dy = (np.roll(f, -1, 1) - f) / hy
        dz = (np.roll(f, -1, 0) - f) / hz
        if p < prm["nac"]:
            e += 0.5 * ep * (kk[0] * dx ** 2 + kk[1] * dy ** 2 + kk[2] * dz ** 2).sum()
            s2 += f ** 2
            s3 += f ** 3
        else:
            e += (0.5 * ep * (dx ** 2 + dy ** 2 + dz ** 2) + 0.25 * (f ** 2 - 1) ** 2).sum()
    e += (0.5 * a2 * s2 - a3 / 3.0 * s3 + 0.25 * a4 * s2 ** 2).sum()
    return e * hx * hy * hz


def test_acceptance_energy_scheme_and_elastic_terms():
    """The acceptance energy of an accepted state: the scheme energy plus
    ES/2 h^3 sum_p eta_p felas_p, against a numpy restatement and felas from
    etd_k_elastic on a second context holding the same fields (FP64); with
    energy_kind 0 the check_soln_new energy replaces the scheme energy."""
    say("acceptance energy: scheme + elastic")
    prm = small_params(n=16, nac=3, nch=1, elastic=1, dt=0.01, h=(0.9, 1.0, 1.1), kk=(1.0, 1.2, 0.8))
    f = noise_fields(prm, amp=0.2)
    for kind in (1, 0):
        g = gpu(prm, f)
        st, log = g.run_adaptive(log_cap=4, t_end=0.01, dt0=0.01, tol=1e9, energy_test=0, energy_kind=kind)
        assert st["accepted"] == 1
        u1 = [g.get_field(v) for v in range(4)]
        h = gpu(prm, u1)
        h.transfer()
        fel = h.elastic()
        sd, _ = h.diag()
        gg = prm["g"]
        inner = (slice(gg, -gg),) * 3
        e_el = 0.5 * 0.9 * 1.0 * 1.1 * sum(float((u1[p][inner] * fel[p][inner]).sum()) for p in range(3))
        e_base = scheme_energy(prm, u1) if kind == 1 else sd[0, 3] + sd[3, 3]
        say(f"  kind {kind}: E = {log[0, 3]:.12e}  (base {e_base:.6e} + elastic {e_el:.6e})")
        assert abs(log[0, 3] - (e_base + e_el)) < 1e-10 * abs(e_base + e_el)


def test_scheme_energy_decreases_along_small_steps():
    """The scheme energy is a Lyapunov function of the discrete dynamics: it
    does not increase over small steps, even from rough states where the
    check_soln_new diagnostic is a different quadratic form."""
    say("scheme energy monotone")
    prm = dict(small_params(n=20, nac=4, nch=1, elastic=1, dt=0.002, KE=3.2, landau=(0.2, 3.0, 3.0)),
               eps0=pyetd.workloads.eigenstrains(4, 0.03, 3))
    f = noise_fields(prm, amp=0.25, ch_amp=0.2)
    g = gpu(prm, f)
    st, log = g.run_adaptive(log_cap=100, t_end=0.2, dt0=0.002, tol=1e9, grow=1.0, shrink=1.0, energy_test=0)
    E = log[:, 3]
    say(f"  E {E[0]:.6e} -> {E[-1]:.6e} over {len(E)} steps, max rise {np.diff(E).max():.3e}")
    assert (np.diff(E) <= 1e-12 * np.abs(E[:-1])).all()


def test_mlp_forward_and_gradient_match_oracle():
    say("mlp forward / grad vs numpy oracle")
    p = M.init(11)
    p[-1] = 0.25
    rng = np.random.default_rng(2)
    x = rng.standard_normal((37, M.NF))
    t = rng.standard_normal(37)
    m = pyetd.Mlp()
    m.set(p.astype(np.float32), np.zeros(16), np.ones(16), 0.0, 1.0)
    p32 = p.astype(np.float32).astype(np.float64)
    x32 = x.astype(np.float32).astype(np.float64)
    t32 = t.astype(np.float32).astype(np.float64)
    y_ref, _ = M.forward(p32, x32)
    y = m.forward(x)
    assert np.abs(y - y_ref).max() < 1e-5 * max(1.0, np.abs(y_ref).max())
    loss_ref, g_ref = M.loss_grad(p32, x32, t32)
    loss, g = m.loss_grad(x, t)
    assert abs(loss - loss_ref) < 1e-5 * loss_ref
    assert np.abs(g - g_ref).max() < 1e-5 * np.abs(g_ref).max()
    m.close()


def test_mlp_training_is_deterministic_and_fits():
    say("mlp training")
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2048, 16)).astype(np.float32)
    y = (np.sin(x[:, 0]) + 0.5 * x[:, 1] ** 2 - 0.3 * x[:, 2] * x[:, 3]).astype(np.float32)
    runs = []
    for _ in range(2):
        m = pyetd.Mlp(seed=3)
        el = m.train(x, y, epochs=40, batch=128, lr=3e-3, seed=9)
        runs.append((el, m.get()[0]))
        m.close()
    say(f"  loss {runs[0][0][0]:.4f} -> {runs[0][0][-1]:.4f}")
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][0], runs[1][0])
    assert runs[0][0][-1] < 0.1 * runs[0][0][0]


def test_learned_controller_keeps_guarantees_and_reports_gain():
    say("learned controller on 32^3 x 4")
    base = dict(small_params(n=36, nac=4, nch=0, elastic=1, dt=0.005, KE=3.2, landau=(0.2, 3.0, 3.0)),
                C=pyetd.workloads.cubic_voigt(134.0, 110.0, 36.0))
    kw = dict(t_end=3.0, tol=1e-4, dt0=0.005, dt_min=1e-5, dt_max=0.5)
    X, Y = [], []
    for seed in range(6):
        prm = dict(base, eps0=pyetd.workloads.eigenstrains(4, 0.03, seed), seed=seed)
        st, log = CT.plain_run(prm, **kw)
        x, y = CT.targets_from_log(log, kw["tol"], kw["dt_min"], kw["dt_max"])
        X.append(x)
        Y.append(y)
    X, Y = np.concatenate(X), np.concatenate(Y)
    m = pyetd.Mlp(seed=5)
    el = CT.fit(m, X, Y, epochs=300, batch=64, lr=2e-3)
    say(f"  {len(Y)} pairs, loss {el[0]:.3f} -> {el[-1]:.3f}")
    prm = dict(base, eps0=pyetd.workloads.eigenstrains(4, 0.03, 99), seed=99)  # held out
    st_p, log_p = CT.plain_run(prm, **kw)
    g = pyetd.Etd.from_dict(prm)
    for v, f in enumerate(pyetd.workloads.initial_fields(prm)):
        g.set_field(v, f)
    g.set_controller(m)
    st_l, log_l = g.run_adaptive(log_cap=100000, controller="learned", **kw)
    say(f"  plain   {st_p}")
    say(f"  learned {st_l}")
    say(f"  step evaluations: plain {st_p['step_evals']}, learned {st_l['step_evals']} "
        f"(gain {st_p['step_evals'] / st_l['step_evals']:.2f}x)")
    for st, log in ((st_p, log_p), (st_l, log_l)):
        assert abs(st["t"] - kw["t_end"]) < 1e-12
        assert (log[:, 2] <= kw["tol"]).all() or (log[:, 1] <= kw["dt_min"] * (1 + 1e-12)).any()
        E = log[:, 3]
        assert (np.diff(E) <= 1e-12 * np.abs(E[:-1])).all()
    assert st_l["learned_proposals"] > 0
    g.close()
    m.close()


@pytest.mark.parametrize("precision", [0, 2])
def test_nan_state_is_an_error_and_is_left_untouched(precision):
    """A NaN cell in the state (set from the host: a data value, not a device
    fault) makes the run stop with ETD_ERR_NONFINITE before any step; nothing
    is accepted and the fields are what was set."""
    say(f"adaptive: NaN state precision={precision}")
    prm = dict(small_params(n=16, nac=3, nch=1, elastic=1, dt=0.01), precision=precision)
    f = noise_fields(prm, amp=0.05)
    f[1] = f[1].copy()
    f[1][7, 8, 9] = np.nan
    g = gpu(prm, f)
    with pytest.raises(pyetd.EtdError) as ei:
        g.run_adaptive(log_cap=8, t_end=0.1, dt0=0.01, dt_min=1e-4, tol=1e-4)
    assert ei.value.code == pyetd.ERR_NONFINITE and "not finite" in str(ei.value)
    assert ei.value.stats["accepted"] == 0 and g.time == 0.0
    for v in range(4):
        a, b = g.get_field(v), f[v].astype(np.float32).astype(np.float64) if precision else f[v]
        assert np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
    g.close()


def test_overflowing_candidates_are_rejected_down_to_dt_min_then_an_error():
    """One cell inflated to 1e75: the state's energy is finite (1e300), every
    candidate overflows (eta ~ -dt a4 eta^3 ~ 1e222, eta^4 = inf), so each is
    rejected with the factor shrink until dt_min, where the run stops with
    ETD_ERR_NONFINITE (before the fix a step at dt_min was accepted whenever
    the error estimate was finite) and the state is restored; the CPU
    restatement makes the same decisions."""
    say("adaptive: overflowing candidates")
    prm = small_params(n=16, nac=3, nch=1, elastic=0, dt=0.01)
    f = noise_fields(prm, amp=0.05)
    f[0] = f[0].copy()
    f[0][8, 8, 8] = 1e75
    kw = dict(t_end=0.1, dt0=0.01, dt_min=1e-4, dt_max=0.1, tol=1e-4)
    o = O.Oracle(prm)
    for v, x in enumerate(f):
        o.set_field(v, x)
    with pytest.raises(AD.NonFinite):
        AD.run(o, **kw)
    g = gpu(prm, f)
    with pytest.raises(pyetd.EtdError) as ei:
        g.run_adaptive(log_cap=8, energy_kind=0, energy_elastic=0, **kw)
    st = ei.value.stats
    say(f"  {st}: {ei.value}")
    assert ei.value.code == pyetd.ERR_NONFINITE and "dt_min" in str(ei.value)
    # dt 1e-2 -> 2e-3 -> 4e-4 -> 1e-4 (clamped): four attempts, all rejected
    assert st["accepted"] == 0 and st["rejected_err"] + st["rejected_energy"] == 3 and st["step_evals"] == 12
    for v in range(4):
        assert np.array_equal(g.get_field(v), f[v])
    g.close()


def test_inflated_field_is_cured_by_smaller_steps():
    """Fields inflated far past the Landau minima (|eta| up to 2.5): the first
    attempts fail the error test, smaller steps are accepted, every accepted
    state is finite and within tolerance, and the energy falls."""
    say("adaptive: inflated field")
    prm = dict(small_params(n=20, nac=3, nch=1, elastic=1, dt=0.05, KE=3.2, landau=(0.2, 3.0, 3.0)),
               eps0=pyetd.workloads.eigenstrains(3, 0.03, 1))
    # periodic ghost images: the ghost fill copies edge and corner ghosts from
    # the pre-fill values of other ghost cells (trans.c:154-251), so random
    # ghosts would make one fill (full step) and two fills (half steps)
    # disagree there by O(1) at any dt
    gg, n = prm["g"], prm["n"]
    idx = (np.arange(n) - gg) % (n - 2 * gg)
    f = [x[gg:-gg, gg:-gg, gg:-gg][np.ix_(idx, idx, idx)] for x in noise_fields(prm, amp=2.5, ch_amp=0.4)]
    g = gpu(prm, f)
    st, log = g.run_adaptive(log_cap=10000, t_end=0.5, dt0=0.05, dt_min=1e-6, dt_max=0.5, tol=1e-4)
    say(f"  {st}")
    assert st["rejected_err"] > 0 and abs(st["t"] - 0.5) < 1e-12
    assert (log[:, 2] <= 1e-4).all()
    E = log[:, 3]
    # accepted when E_new <= E + 1e-12 x (sum of the energy's absolute terms)
    assert np.isfinite(E).all() and (np.diff(E) <= 1e-10 * np.abs(E).max()).all() and E[-1] < E[0]
    assert all(np.isfinite(g.get_field(v)).all() for v in range(4))
    g.close()

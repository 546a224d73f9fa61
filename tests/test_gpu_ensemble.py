"""GPU tests of batched ensembles (one hipGraph over many contexts of uneven
sizes) and the kernel-variant autotuner, through the C ABI.

A graph replay must equal stepping each member on its own, bit for bit: the
graph holds the same kernels with the same arguments.  Each replay is checked
against the eager result, with eager work (diagnostics, field reads) and
synchronisations between the replays."""
import sys

import numpy as np
import pytest

import pyetd
from conftest import noise_fields, small_params

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


MEMBERS = [
    dict(n=36, nac=4, nch=0, etd2=0, elastic=1, aniso=0, precision=0),   # N = 32, FP64
    dict(n=20, nac=2, nch=1, etd2=1, elastic=0, aniso=1, precision=0),   # ETD2 + CH
    dict(n=52, nac=4, nch=0, etd2=0, elastic=1, aniso=0, precision=1),   # N = 48, mixed BF16
    dict(n=68, nac=4, nch=1, etd2=0, elastic=1, aniso=0, precision=2),   # N = 64, mixed FP16 + CH mean
]


def build(spec, seed, carry=True):
    q = dict(spec)
    prec = q.pop("precision")
    prm = dict(small_params(dt=0.01, **q), precision=prec)
    g = pyetd.Etd.from_dict(prm, carry=carry)
    for v, f in enumerate(noise_fields(prm, seed=seed, amp=0.2)):
        g.set_field(v, f)
    return g


def test_ensemble_replays_equal_solo_stepping():
    say("ensemble of 4 uneven members, 2 replays of 3 steps")
    ens_m = [build(s, 10 + i) for i, s in enumerate(MEMBERS)]
    # the graph replays steps that transform u; so do the solo runs
    solo = [build(s, 10 + i, carry=False) for i, s in enumerate(MEMBERS)]
    K = 3  # nac = 4, ETD1: the felas ring (6 slots) repeats every 3 steps
    E = pyetd.Ensemble(ens_m, K)
    for rep in range(2):
        E.run(1)
        for g in solo:
            g.step(K)
        for i, (a, b) in enumerate(zip(ens_m, solo)):
            for v in range(a.V):
                assert np.array_equal(a.get_field(v), b.get_field(v)), (rep, i, v)
            assert a.time == b.time
            a.diag()  # eager work between replays
    E.close()
    # the members keep stepping eagerly after the graph, in the same ring phase
    for a, b in zip(ens_m, solo):
        a.step(2)
        b.step(2)
        for v in range(a.V):
            assert np.array_equal(a.get_field(v), b.get_field(v))


def test_ensemble_rejects_misaligned_step_count():
    say("ensemble misaligned")
    g = build(MEMBERS[0], 1)
    with pytest.raises(pyetd.EtdError, match="multiple of 3"):
        pyetd.Ensemble([g], 2)


def test_autotune_picks_a_valid_choice():
    say("autotune")
    for spec in (MEMBERS[0], MEMBERS[2]):
        q = dict(spec)
        prec = q.pop("precision")
        prm = dict(small_params(dt=0.01, **q), precision=prec)
        t = pyetd.autotune(prm, steps=2)
        say(f"  n={prm['n']} precision={prec}: {t.as_dict()}")
        assert t.ms_per_step > 0 and t.fft_variant in (0, 1)
        a, b = build(spec, 3), build(spec, 3)
        a.set_tuning(t)
        a.step(2)
        b.step(2)
        tol = 1e-12 if prec == 0 else 1e-6
        for v in range(a.V):
            assert np.abs(a.get_field(v) - b.get_field(v)).max() < tol

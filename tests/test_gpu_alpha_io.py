"""The beta -> alpha driver's output and restart formats on the GPU, through
the C ABI, against the CPU restatement (oracle/formats.py) and the byte
layouts of source_beta_to_alpha/src/vars.c:967-1094:

- the binary checkpoint: int irun + n^3 doubles per field, read back exactly;
- the uint16 dumps (eta + 0.1) * 30000, every cell, including values that
  wrap and saturate;
- the detector feed: 2x downsample, binarise at 0.1, 8 cells per byte;
- the etd_run --alpha driver: its schedule of dumps, double-buffered
  checkpoints and detector files, and a restart from the checkpoint that
  continues the run bit for bit (ETD1 without elasticity: the step depends
  only on the fields)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import formats as F
import oracle as O
import pyetd
from conftest import ROOT, noise_fields, small_params

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "energy-stability_amd", "etd_run")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def context(prm, fields, work, data):
    p = pyetd.params_from_dict(prm)
    p.work_dir = (work.rstrip("/") + "/").encode()
    p.data_dir = (data.rstrip("/") + "/").encode()
    Vm, Vim, D = O.eigbasis(prm["n"], 3)
    g = pyetd.Etd(p, Vm, Vim, D)
    for v, f in enumerate(fields):
        g.set_field(v, f)
    return g


@pytest.mark.parametrize("precision", [0, 2])
def test_checkpoint_round_trip_and_layout(tmp_path, precision):
    say(f"checkpoint precision={precision}")
    os.makedirs(tmp_path / "out" / "00")
    prm = dict(small_params(n=20, nac=3, nch=1), precision=precision)
    f = noise_fields(prm, amp=0.4)
    g = context(prm, f, str(tmp_path / "out"), str(tmp_path))
    before = [g.get_field(v) for v in range(4)]
    g.write_checkpoint(7, 1)
    for v in range(4):
        name = f"eta{v}_chk1_000000.dat" if v < 3 else "c0_chk1_000000.dat"
        raw = open(tmp_path / "out" / "00" / name, "rb").read()
        assert len(raw) == 4 + 8 * 20 ** 3
        assert np.frombuffer(raw[:4], np.int32)[0] == 7
        assert np.array_equal(np.frombuffer(raw[4:], np.float64).reshape(20, 20, 20), before[v])
    g.step(2)
    assert g.read_checkpoint(1) == 7
    for v in range(4):
        assert np.array_equal(g.get_field(v), before[v])
    with pytest.raises(pyetd.EtdError):
        g.read_checkpoint(0)  # never written
    g.close()


@pytest.mark.parametrize("precision", [0, 2])
def test_uint16_dumps_match_restatement(tmp_path, precision):
    say(f"uint16 dumps precision={precision}")
    os.makedirs(tmp_path / "data" / "00")
    prm = dict(small_params(n=16, nac=2, nch=1), precision=precision)
    rng = np.random.default_rng(5)
    f = [rng.uniform(-0.5, 2.5, size=(16, 16, 16)) for _ in range(3)]  # wraps above 2.08, saturates below -0.1
    g = context(prm, f, str(tmp_path), str(tmp_path / "data"))
    g.write_field2b(12)
    for v in range(2):
        raw = np.fromfile(tmp_path / "data" / "00" / f"eta{v}_000012_000000.dat", np.uint16)
        ref = F.field2b(g.get_field(v)).reshape(-1)
        assert raw.size == 16 ** 3 and np.array_equal(raw, ref)
    assert not os.path.exists(tmp_path / "data" / "00" / "eta2_000012_000000.dat")  # CH fields are not dumped
    g.close()


@pytest.mark.parametrize("precision,n", [(0, 36), (2, 36), (0, 20)])
def test_detector_feed_matches_restatement(tmp_path, precision, n):
    say(f"detector feed precision={precision} n={n}")
    prm = dict(small_params(n=n, nac=2), precision=precision)
    f = noise_fields(prm, amp=0.25)
    f = [x + 0.08 for x in f]  # around the 0.1 threshold
    g = context(prm, f, str(tmp_path), str(tmp_path))
    for v in range(2):
        bits, binary = g.detector_feed(v)
        rb, rbin = F.detector(g.get_field(v))
        assert np.array_equal(bits, rb) and np.array_equal(binary, rbin)
        assert 0.2 < binary.mean() < 0.8
        bits2, none = g.detector_feed(v, want_binary=False)
        assert none is None and np.array_equal(bits2, rb)
    g.close()


def alpha_case(d, checkpoint, chk, t_total, elastic=0, n=36):
    ini = (f" {n}, {n}, {n}\n 4, 4, 4\n 2\n ./out/\n ./data/\n 0.01, {t_total!r}\n 0\n 1\n 3, 0\n {elastic}\n 0\n"
           f" 0, 0, 0\n {n - 4}, {n - 4}, {n - 4}\n 1.0, 1.0, 1.0\n {checkpoint}, 100, {chk}\n 50\n 16\n")
    open(os.path.join(d, "input9.ini"), "w").write(ini)
    open(os.path.join(d, "landau_input.txt"), "w").write("1.0 3.2 0.2 3.0 3.0\n")
    for sub in ("out/00", "data/00", f"downsample_output_{n}_111"):
        os.makedirs(os.path.join(d, sub), exist_ok=True)


def run_alpha(d):
    # ETD_CARRY=0: a restarted run reproduces an unbroken one bit for bit
    # (with the carried spectrum the two agree to roundoff)
    env = dict(os.environ, ETD_SEED="5", ETD_DETECT_EVERY="100", ETD_CARRY="0")
    r = subprocess.run([DRIVER, "--alpha", "9", "."], cwd=d, env=env, capture_output=True, text=True, timeout=300)
    say(r.stdout[-300:])
    assert r.returncode == 0, r.stderr


def read_chk(path):
    raw = open(path, "rb").read()
    return int(np.frombuffer(raw[:4], np.int32)[0]), np.frombuffer(raw[4:], np.float64).reshape(36, 36, 36)


def test_alpha_driver_schedule_and_restart(tmp_path):
    """300 steps in one run against 200 steps, then a restart from the
    checkpoint of step 200 for 100 more: the same final checkpoint bytes."""
    a = str(tmp_path / "a")
    os.makedirs(a)
    alpha_case(a, 1, 0, 3.0)  # 300 steps of dt 0.01
    say("alpha driver: 300 steps")
    run_alpha(a)
    b = str(tmp_path / "b")
    os.makedirs(b)
    alpha_case(b, 1, 0, 2.0)
    say("alpha driver: 200 steps")
    run_alpha(b)
    # dumps every 50 steps (irun 1..4), checkpoints every 100 (buffers 0, 1)
    for irun in range(1, 5):
        for v in range(3):
            assert os.path.getsize(os.path.join(b, "data", "00", f"eta{v}_{irun:06d}_000000.dat")) == 2 * 36 ** 3
    i0, u0 = read_chk(os.path.join(b, "out", "00", "eta0_chk0_000000.dat"))
    i1, u1 = read_chk(os.path.join(b, "out", "00", "eta0_chk1_000000.dat"))
    assert (i0, i1) == (2, 4)
    # the last dump and the detector file of iteration 200 are those of the step-200 state
    dump = np.fromfile(os.path.join(b, "data", "00", "eta0_000004_000000.dat"), np.uint16)
    assert np.array_equal(dump, F.field2b(u1).reshape(-1))
    feed = open(os.path.join(b, f"downsample_output_36_111", "eta0_200_0_0_0_ds.txt")).read().split()
    assert [int(x) for x in feed] == list(F.detector(u1)[0])
    assert os.path.exists(os.path.join(b, f"downsample_output_36_111", "eta2_100_0_0_0_ds.txt"))
    # restart from buffer 1 (step 200) for 100 steps: checkpoint 2 continues in buffer 0
    alpha_case(b, 2, 1, 1.0)
    say("alpha driver: restart from buffer 1, 100 steps")
    run_alpha(b)
    for v in range(3):
        ia, ua = read_chk(os.path.join(a, "out", "00", f"eta{v}_chk0_000000.dat"))  # run a: step 300 -> buffer 0
        ib, ub = read_chk(os.path.join(b, "out", "00", f"eta{v}_chk0_000000.dat"))
        assert ia == 6 and ib == 6 and np.array_equal(ua, ub), v


def test_alpha_driver_runs_the_alpha_physics(tmp_path):
    """etd_run --alpha with the driver's own physics: CALPHAD driving force,
    its compact gradient stencil (ANISOTROPIC = 1), LE = 2, KE = 3,
    epn2 = 0.048, and a grain mask file (ETD_ALPHA_MASK); the step-100
    checkpoint against the oracle stepping the same seeded state."""
    n = 20
    d = str(tmp_path)
    ini = (f" {n}, {n}, {n}\n 1, 1, 1\n 2\n ./out/\n ./data/\n 0.01, 1.0\n 0\n 1\n 2, 0\n 0\n 1\n"
           f" 0, 0, 0\n {n - 4}, {n - 4}, {n - 4}\n 1.0, 1.0, 1.0\n 1, 100, 0\n 1000\n 16\n")
    open(os.path.join(d, "input7.ini"), "w").write(ini)
    rng = np.random.default_rng(11)
    lam_in = rng.uniform(-0.2, 0.2, size=(2, 3, 3))
    open(os.path.join(d, "anisotropic_input.txt"), "w").write(" ".join(repr(float(x)) for x in lam_in.reshape(-1)) + "\n")
    mask = (rng.uniform(size=(n, n, n)) > 0.3).astype(np.int32)
    mask.tofile(os.path.join(d, "grains.bin"))
    for sub in ("out/00", "data/00", f"downsample_output_{n}_111"):
        os.makedirs(os.path.join(d, sub), exist_ok=True)
    env = dict(os.environ, ETD_SEED="5", ETD_ALPHA_MASK=os.path.join(d, "grains.bin"))
    say("alpha driver, alpha physics: 100 steps")
    r = subprocess.run([DRIVER, "--alpha", "7", "."], cwd=d, env=env, capture_output=True, text=True, timeout=300)
    say(r.stdout[-200:])
    assert r.returncode == 0, r.stderr
    # the same initial state: etd_init_noise(seed + 7919 v), AC mean 0 amplitude 0.01
    prm = dict(small_params(n=n, nac=2), free_energy=1, LE=2.0, KE=3.0, epn2=0.048, dt=0.01,
               lam=np.eye(3)[None, :, :] - lam_in, mask=mask)
    prm["aniso"] = 2
    g = pyetd.Etd.from_dict(prm)
    f0 = []
    for v in range(2):
        g.init_noise(v, 5 + 7919 * v, 0.0, 0.01)
        f0.append(g.get_field(v))
    g.close()
    o = O.Oracle(prm)
    for v in range(2):
        o.set_field(v, f0[v])
    o.step(100)
    for v in range(2):
        raw = open(os.path.join(d, "out", "00", f"eta{v}_chk0_000000.dat"), "rb").read()
        got = np.frombuffer(raw[4:], np.float64).reshape(n, n, n)
        err = float(np.abs(got - o.get_field(v)).max())
        say(f"  field {v}: max |d| vs oracle after 100 steps = {err:.2e}")
        assert err < 1e-9

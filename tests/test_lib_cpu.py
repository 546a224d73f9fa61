"""Host-side checks of libetd that need no GPU: the library loads, exports every
entry point declared in include/etd.h, parses the reference's own input files,
generates the same eigenbasis as the reference's shipped files, and fails with
an error code (not a crash, not a CPU fallback) when no device is present."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyetd
from conftest import GOLDEN, ROOT


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "etd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(etd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = pyetd.lib()
    syms = header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s


def test_parse_omega_ini_fixture():
    p = pyetd.default_params()
    rc = pyetd.lib().etd_parse_ini(os.path.join(GOLDEN, "omega", "input64.ini").encode(), C.byref(p))
    assert rc == 0
    assert (p.nx, p.ny, p.nz) == (68, 68, 68)
    assert (p.px, p.py, p.pz) == (2, 2, 2)
    assert p.nghost == 2
    assert p.work_dir == b"./output/" and p.data_dir == b"./data/"
    assert p.dt == 1e-6 and p.t_total == 0.0066
    assert p.epn2 == 0.6955
    assert p.periodic == 1 and (p.nac, p.nch) == (4, 0)
    assert (p.restart, p.restart_iter) == (0, 0) and p.restart_dir == b"./restart/"
    assert (p.xmax, p.ymax, p.zmax) == (136.0, 136.0, 136.0)
    assert (p.etd2, p.elastic, p.anisotropic) == (1, 1, 1)
    assert (p.kkx, p.kky, p.kkz) == (2.5, 2.5, 2.5)
    assert p.variable_mobility == 0
    # init_para: gnx = 68*2 - 2*2*2 = 128
    assert p.hx == 136.0 / 128


def test_parse_alpha_ini_fixture():
    p = pyetd.default_params()
    rc = pyetd.lib().etd_parse_ini_alpha(os.path.join(GOLDEN, "alpha", "input192.ini").encode(), C.byref(p))
    assert rc == 0
    assert (p.nx, p.px, p.nghost) == (260, 4, 2)
    assert (p.dt, p.t_total) == (2.0, 10000.0)
    # the driver's ANISOTROPIC = 1 is its compact stencil (anisotropic = 2 here)
    assert (p.nac, p.nch, p.elastic, p.anisotropic) == (6, 2, 1, 2)
    assert p.hx == 1.0  # 1024 / (260*4 - 16)
    assert p.epn2 == 0.048 and p.etd2 == 0
    # its physics: CALPHAD driving force, LE = 2, KE = 3 (vars.c:114-127)
    assert (p.free_energy, p.LE, p.KE) == (1, 2.0, 3.0)


def test_parse_elastic_and_anisotropic_fixtures():
    p = pyetd.parse_inputs(os.path.join(GOLDEN, "omega"), "64")
    assert p.elastic_scale == 5.9022
    C36 = np.array(p.C[:])
    assert C36[0] == 102.12 and C36[1] == 79.63 and C36[21] == 34.25 and C36[35] == 34.25
    eps = np.array(p.eps0[: 4 * 9]).reshape(4, 3, 3)
    assert eps[0, 0, 0] == -0.0016 and eps[0, 0, 1] == 0.0054
    assert eps[1, 1, 2] == 0.0054 and eps[3, 2, 0] == -0.0054
    lam = np.array(p.lam[: 4 * 9]).reshape(4, 3, 3)
    # Lambda = diag(kk) - lambda_in (anisotropic.c:664-673), kk = 2.5
    assert lam[0, 0, 0] == 1.5 and lam[0, 0, 2] == -1.0 and lam[1, 0, 2] == 1.0 and lam[0, 2, 2] == 1.5


def test_missing_inputs_are_errors(tmp_path):
    p = pyetd.default_params()
    assert pyetd.lib().etd_parse_ini(str(tmp_path / "input0.ini").encode(), C.byref(p)) == -2
    (tmp_path / "input1.ini").write_text("68, 68\n")
    assert pyetd.lib().etd_parse_ini(str(tmp_path / "input1.ini").encode(), C.byref(p)) == -2
    assert pyetd.lib().etd_parse_elastic(str(tmp_path / "elastic_input.txt").encode(), C.byref(p)) == -2


@pytest.mark.parametrize("bcid", [0, 1, 2, 3])
def test_product_eigbasis_matches_reference_files(bcid):
    z = np.load(os.path.join(GOLDEN, "eig260.npz"))
    d = z[f"d{bcid}"]
    V, Vi, D = pyetd.eigbasis(260, bcid)
    assert np.abs(np.sort(D) - np.sort(d)).max() < 5e-14
    rows = z["rows"]
    vrow = z[f"vrow{bcid}"]
    for c in range(0, 260, 7):
        cc = int(np.argmin(np.abs(D - d[c])))
        mine = V[rows, cc]
        scale = np.dot(vrow[:, c], vrow[:, c]) / np.dot(mine, vrow[:, c])
        assert np.abs(mine * scale - vrow[:, c]).max() < 1e-10
    T = V @ np.diag(D) @ Vi
    assert abs(T[5, 5] + 2.0) < 1e-12 and abs(T[5, 6] - 1.0) < 1e-12


def test_read_matrices_file_format(tmp_path):
    """d/v/vi<id><n>.dat: row-major text in the file, column-major in memory."""
    n = 6
    V, Vi, D = pyetd.eigbasis(n, 3)
    np.savetxt(tmp_path / f"d3{n}.dat", D, fmt="%.16e")
    np.savetxt(tmp_path / f"v3{n}.dat", V, fmt="%.16e")
    np.savetxt(tmp_path / f"vi3{n}.dat", Vi, fmt="%.16e")
    a, b, c = np.zeros(n * n), np.zeros(n * n), np.zeros(n)
    _p = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
    assert pyetd.lib().etd_read_matrices(str(tmp_path).encode(), n, 3, _p(a), _p(b), _p(c)) == 0
    assert np.array_equal(a.reshape(n, n).T, V)
    assert np.array_equal(b.reshape(n, n).T, Vi)
    assert np.array_equal(c, D)
    assert pyetd.lib().etd_read_matrices(str(tmp_path).encode(), n, 2, _p(a), _p(b), _p(c)) == -2


def test_unsupported_configurations_are_rejected():
    """Decomposed runs outside what is built (mixed precision, physical outer
    sides, a non-cubic grid of boxes under elasticity), non-cubes and variable
    mobility without its three CH fields return ETD_ERR_UNSUPPORTED before
    any device work."""
    cases = (dict(px=2, precision=2), dict(px=2, periodic=0), dict(px=2, elastic=1), dict(variable_mobility=1),
             dict(ny=14))
    for over in cases:
        p = pyetd.params_from_dict(dict(n=12, nac=1, nch=0, dt=0.01, epn2=1.0))
        for field, val in over.items():
            setattr(p, field, val)
        ctx = C.c_void_p()
        rc = pyetd.lib().etd_create(C.byref(p), None, None, None, C.byref(ctx))
        assert rc == -3, over


def test_precision_flag_is_checked_before_device_work():
    """Mixed precision needs an edge divisible by 4 (16-byte FP32 copies);
    unknown precision values are argument errors."""
    p = pyetd.params_from_dict(dict(n=14, nac=1, nch=0, dt=0.01, epn2=1.0, precision=1))
    ctx = C.c_void_p()
    assert pyetd.lib().etd_create(C.byref(p), None, None, None, C.byref(ctx)) == -3
    p = pyetd.params_from_dict(dict(n=16, nac=1, nch=0, dt=0.01, epn2=1.0, precision=7))
    assert pyetd.lib().etd_create(C.byref(p), None, None, None, C.byref(ctx)) == -1


def test_seeded_nuclei_initial_state():
    """The coarsening workload's nuclei: deterministic for a seed, periodic
    ghosts, values within [-a, 1], and ordered volume close to the sum of the
    nuclei's spheres (overlaps only reduce it)."""
    from pyetd import workloads as W
    prm = W.scaled(W.beta_omega_256_coarsening(), 68)
    a = W.initial_fields(prm)
    b = W.initial_fields(prm)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    g, N = prm["g"], 64
    for x in a:
        assert np.array_equal(x[:g], x[N:N + g]) and np.array_equal(x[:, :, -g:], x[:, :, g:2 * g])
        assert x.min() >= -0.01 and x.max() <= 1.0
    inner = (slice(g, -g),) * 3
    frac = sum(float((x[inner] > 0.5).mean()) for x in a)
    count = round(prm["nuclei_per_cell"] * N ** 3)
    r3 = (3.0 ** 4 - 6.0 ** 4) / (4 * (3.0 - 6.0))  # E[r^3] for r ~ U[3, 6]
    expect = count * 4.0 / 3.0 * np.pi * r3 / N ** 3
    assert 0.6 * expect < frac <= 1.05 * expect

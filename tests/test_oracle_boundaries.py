"""Pins the oracle's physical (non-periodic) boundaries on one box
(periodic = 0, procs = 1,1,1), which the reference reaches through
init_para (vars.c:58-157: left < 0 gives ix2 = 0, ix3 = nx - nghost,
gnx = nx), read_matrices (vars.c:488-582: boundary id 0 on every axis) and
the left < 0 branches of its stencils:

- one ETD1 step with N == 0 is exp(dt A) u exactly, A the Kronecker sum of
  the reflecting second-difference matrix (no outside values, no ghost fill);
- omega's F2 and CH Laplacian mirror the low side and read 0 past the high
  side (anisotropic.c:177-216, calc.c): hand-computed known answers;
- the compact stencil mirrors both sides (anisotropic_hip.c:546-591);
- check_soln_new runs over [0, n - g) (field.c:457-527 with ix2 = 0);
- the elastic force keeps 0 outside the interior (no exchange)."""
import numpy as np
from scipy.linalg import expm

import oracle as O
from conftest import small_params


def phys(LE=None, **kw):
    prm = small_params(**kw)
    prm["periodic"] = 0
    if LE is not None:
        prm["LE"] = LE
    return prm


def test_physical_etd1_linear_step_is_exact_expm():
    n = 7
    prm = phys(n=n, nac=1, landau=(0.0, 0.0, 0.0), KE=0.0, kk=(1.3, 0.8, 1.1), epn2=0.7, h=(1.0, 0.9, 1.2),
               dt=0.05)
    o = O.Oracle(prm)
    rng = np.random.default_rng(12)
    u0 = rng.standard_normal((n, n, n))
    o.set_field(0, u0)
    o.step(1)
    T = np.diag(-2.0 * np.ones(n)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    T[0, 1] = 2.0  # reflecting ends (boundary id 0)
    T[n - 1, n - 2] = 2.0
    I = np.eye(n)
    hx, hy, hz = prm["h"]
    kx, ky, kz = prm["kk"]
    A = kx * np.kron(I, np.kron(I, T)) / hx ** 2 + ky * np.kron(I, np.kron(T, I)) / hy ** 2 + kz * np.kron(T, np.kron(I, I)) / hz ** 2
    A *= prm["LE"] * prm["epn2"]
    u1 = expm(prm["dt"] * A) @ u0.reshape(-1)
    assert np.abs(o.get_field(0).reshape(-1) - u1).max() < 1e-12


def test_physical_f2_neighbour_rule_known_answers():
    """u = i along x, Lambda = I, h = 1: g_x = (u[i+1] - u[i-1]) / 2 with u[-1] ->
    u[1] (mirror) and u[n] -> 0, F2 = (g[i+1] - g[i-1]) / 2 with the same rule:
    n = 6 gives g = (0, 1, 1, 1, 1, -2), F2 = (0, 1/2, 0, 0, -3/2, -1/2)."""
    n = 6
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    u = i.astype(float)
    res = []
    for aniso in (0, 1):
        prm = dict(phys(n=n, nac=1, LE=1.0, epn2=1.0, KE=0.0, landau=(0.0, 0.0, 0.0)), aniso=aniso, lam=np.eye(3)[None])
        o = O.Oracle(prm)
        o.set_field(0, u)
        res.append(o.nonlinear(0))
    f2 = -(res[1] - res[0])
    want = np.array([0.0, 0.5, 0.0, 0.0, -1.5, -0.5])[None, None, :] + 0 * u
    # along y (and z) the field w = i is constant: g_y = (w[j-1] - w[j+1]) / 2 is 0 except
    # w / 2 at j = n - 1 (0 past the box), so F2_y = -w / 4 at j = n - 2 and 0 elsewhere
    want = want - 0.25 * u * (j == n - 2) - 0.25 * u * (k == n - 2)
    assert np.abs(f2 - want).max() < 1e-14


def test_physical_ch_laplacian_rule_and_no_outside_terms():
    """N_c = Lap t per axis: low cell 2 (t[1] - t[0]) / h^2 (mirror), high cell
    (t[n-2] - 2 t[n-1]) / h^2 (0 past the box), no mu_out; t pointwise."""
    n = 6
    rng = np.random.default_rng(13)
    c = 0.2 + 0.1 * rng.standard_normal((n, n, n))
    prm = phys(n=n, nac=1, nch=1, h=(0.9, 1.0, 1.1))
    o = O.Oracle(prm)
    o.set_field(0, np.zeros((n, n, n)))
    o.set_field(1, c)
    N = o.nonlinear(1)
    t = -prm["LCI"] * prm["KCI"] * c + prm["LCI"] * (c * (c * c - 1))
    def axis_term(a, ax, h):
        a = np.moveaxis(a, ax, -1)
        out = np.empty_like(a)
        out[..., 1:-1] = a[..., :-2] + a[..., 2:] - 2 * a[..., 1:-1]
        out[..., 0] = 2 * (a[..., 1] - a[..., 0])
        out[..., -1] = a[..., -2] - 2 * a[..., -1]
        return np.moveaxis(out, -1, ax) / h ** 2
    want = axis_term(t, 2, 0.9) + axis_term(t, 1, 1.0) + axis_term(t, 0, 1.1)
    assert np.abs(N - want).max() < 1e-12 * np.abs(want).max()


def test_physical_compact_stencil_mirrors_both_sides():
    """u = i^2 along x, Lambda = I, h = 1: f_xx = 2 inside, 2 (u[1] - u[0]) = 2 at
    i = 0 and 2 (u[n-2] - u[n-1]) at i = n - 1 (mirror)."""
    n = 6
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    u = (i * i).astype(float)
    res = []
    for aniso in (0, 2):
        o = O.Oracle(dict(phys(n=n, nac=1, LE=1.0, epn2=1.0, KE=0.0, landau=(0.0, 0.0, 0.0)), aniso=aniso))
        o.set_field(0, u)
        res.append(o.nonlinear(0))
    f2 = -(res[1] - res[0])
    want = np.full(n, 2.0)
    want[-1] = 2.0 * ((n - 2) ** 2 - (n - 1) ** 2)
    assert np.abs(f2 - want[None, None, :]).max() < 1e-13


def test_physical_diagnostics_cover_the_low_ghost_cells():
    n, g = 8, 2
    h = (0.9, 1.0, 1.1)
    rng = np.random.default_rng(14)
    u = rng.uniform(-0.5, 0.5, size=(n, n, n))
    u[0, 0, 0] = 3.0  # a low ghost cell of the periodic box
    st = {}
    for periodic in (1, 0):
        prm = dict(small_params(n=n, nac=1, h=h), periodic=periodic)
        o = O.Oracle(prm)
        o.set_field(0, u)
        st[periodic] = o.diag()[0][0]
    assert st[0][0] == 3.0 and st[1][0] < 3.0
    box = u[: n - g, : n - g, : n - g]
    assert st[0][1] == box.min()
    assert abs(st[0][2] - h[0] * h[1] * h[2] * np.sum(box + 1.0)) < 1e-12 * abs(st[0][2])


def test_physical_elastic_force_is_zero_outside_the_interior():
    n, g = 12, 2
    prm = small_params(n=n, nac=2, elastic=1)
    rng = np.random.default_rng(15)
    f = [rng.uniform(-0.3, 0.3, size=(n, n, n)) for _ in range(2)]
    out = {}
    for periodic in (1, 0):
        o = O.Oracle(dict(prm, periodic=periodic))
        for v in range(2):
            o.set_field(v, f[v])
        o.elastic()
        out[periodic] = [o.felas(p) for p in range(2)]
    inner = (slice(g, n - g),) * 3
    for p in range(2):
        assert np.array_equal(out[0][p][inner], out[1][p][inner])
        ghost = np.ones((n, n, n), bool)
        ghost[inner] = False
        assert np.all(out[0][p][ghost] == 0.0) and np.any(out[1][p][ghost] != 0.0)

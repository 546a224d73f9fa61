"""Pins the FP64 CPU oracle against the reference's golden vectors and against
what the mathematics fixes (closed forms, dense matrix exponentials,
brute-force DFTs, fixed points, convergence order)."""
import os

import numpy as np
import pytest
from scipy.linalg import expm

import oracle as O
from conftest import GOLDEN, OMEGA_LANDAU, noise_fields, small_params


# ---------------------------------------------------------------- eigenbasis
@pytest.mark.parametrize("bcid", [0, 1, 2, 3])
def test_eigbasis_matches_reference_files(bcid):
    """Closed-form V, V^-1, D vs source_beta_to_alpha/{d,v,vi}<id>260.dat."""
    z = np.load(os.path.join(GOLDEN, "eig260.npz"))
    d = z[f"d{bcid}"]
    V, Vi, D = O.eigbasis(260, bcid)
    assert np.abs(np.sort(D) - np.sort(d)).max() < 5e-14
    rows, cols = z["rows"], z["cols"]
    vrow, virow = z[f"vrow{bcid}"], z[f"virow{bcid}"]
    for c in range(260):
        cc = int(np.argmin(np.abs(D - d[c])))
        mine = V[rows, cc]
        sgn = np.sign(np.dot(mine, vrow[:, c]))
        scale = np.dot(vrow[:, c], vrow[:, c]) / np.dot(mine, vrow[:, c])  # eigenvector norm is free
        assert sgn != 0
        # the shipped vectors are numerical (error ~ eps*|T|/gap ~ 1e-12 near the
        # spectrum edges); a wrong closed form would miss by O(0.1)
        assert np.abs(mine * scale - vrow[:, c]).max() < 1e-10
    for t, c in enumerate(cols):
        cc = int(np.argmin(np.abs(D - d[c])))
        mine = Vi[cc, :]
        scale = np.dot(virow[t], virow[t]) / np.dot(mine, virow[t])
        assert np.abs(mine * scale - virow[t]).max() < 1e-10


@pytest.mark.parametrize("n", [8, 13, 68])
@pytest.mark.parametrize("bcid", [0, 1, 2, 3])
def test_eigbasis_reconstructs_operator(n, bcid):
    V, Vi, D = O.eigbasis(n, bcid)
    T = np.diag(-2.0 * np.ones(n)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    if bcid in (0, 2):
        T[0, 1] = 2.0
    if bcid in (0, 1):
        T[n - 1, n - 2] = 2.0
    assert np.abs(V @ np.diag(D) @ Vi - T).max() < 1e-13
    assert np.abs(V @ Vi - np.eye(n)).max() < 1e-13


# ---------------------------------------------------------------- mode product
@pytest.mark.parametrize("n", [3, 5, 8])
def test_modeprod_is_kronecker(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    u = rng.standard_normal(n ** 3)
    # dense: contract the fastest axis (a) of u[c, b, a]
    K = np.kron(np.eye(n * n), A)  # acts on a fastest
    dense = (K @ u).reshape(n, n, n)  # [c, b, m]
    out = O.modeprod(A, u, rotate=True).reshape(n, n, n)  # [m, c, b]
    assert np.abs(out - dense.transpose(2, 0, 1)).max() < 1e-13
    out0 = O.modeprod(A, u, rotate=False).reshape(n, n, n)
    assert np.abs(out0 - dense).max() < 1e-13
    # three rotating products = full (A x A x A) Kronecker product, back in xyz
    w = u
    for _ in range(3):
        w = O.modeprod(A, w, rotate=True)
    full = np.kron(A, np.kron(A, A)) @ u
    assert np.abs(w - full).max() < 1e-12


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("N", [2, 3, 5, 6])
def test_fft_brute_force_dft(N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    idx = np.arange(N)
    W = np.exp(-2j * np.pi * np.outer(idx, idx) / N)
    ref = np.einsum("az,by,cx,zyx->abc", W, W, W, x)
    assert np.abs(O.fft3d(x, 1) - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("N", [8, 12, 18, 20])
def test_fft_round_trip_parseval(N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal((N, N, N))
    X = O.fft3d(x, 1)
    assert abs(np.sum(np.abs(X) ** 2) / N ** 3 - np.sum(x ** 2)) < 1e-10 * np.sum(x ** 2)
    back = O.fft3d(X, -1)
    assert np.abs(back - x).max() < 1e-13


# ---------------------------------------------------------------- ghost fill
def test_ghost_fill_keeps_periodic_state():
    n, g = 12, 2
    N = n - 2 * g
    rng = np.random.default_rng(3)
    P = rng.standard_normal((N, N, N))
    idx = (np.arange(n) - g) % N
    u = P[np.ix_(idx, idx, idx)]
    assert np.array_equal(O.ghost_fill(u, g), u)


def test_ghost_fill_face_values_and_interior():
    n, g = 10, 2
    N = n - 2 * g
    rng = np.random.default_rng(4)
    u = rng.standard_normal((n, n, n))
    out = O.ghost_fill(u, g)
    inner = slice(g, n - g)
    assert np.array_equal(out[inner, inner, inner], u[inner, inner, inner])
    # face ghosts (only one ghost coordinate) are the periodic interior images
    for i in range(g):
        assert np.array_equal(out[inner, inner, i], u[inner, inner, i + N])
        assert np.array_equal(out[inner, i, inner], u[inner, i + N, inner])
        assert np.array_equal(out[i, inner, inner], u[i + N, inner, inner])
        assert np.array_equal(out[inner, inner, n - g + i], u[inner, inner, g + i])


# ---------------------------------------------------------------- nonlinear
def test_ac_nonlinear_pointwise_interior():
    prm = small_params(n=10, nac=3)
    o = O.Oracle(prm)
    f = noise_fields(prm, amp=0.5)
    for v in range(3):
        o.set_field(v, f[v])
    a2, a3, a4 = prm["landau"]
    s = sum(fi ** 2 for fi in f)
    for v in range(3):
        N = o.nonlinear(v)
        u = f[v]
        expect = prm["LE"] * prm["KE"] * u - prm["LE"] * (a2 * u - a3 * u ** 2 + a4 * s * u)
        inner = slice(1, 9)
        assert np.abs(N[inner, inner, inner] - expect[inner, inner, inner]).max() < 1e-13


def test_ac_f2_fourier_mode():
    """F2 on cos(k.x) away from the box edges equals
    -sum_ab s_a s_b sig_a sig_b Lambda_ab cos(k.x), s_a = sin(k_a h_a)/h_a,
    sig = (1,-1,-1) (the y and z differences run j-1 minus j+1)."""
    prm = small_params(n=14, nac=1, aniso=1, h=(0.7, 1.1, 0.9))
    prm["landau"] = (0.0, 0.0, 0.0)
    prm["KE"] = 0.0
    lam = prm["lam"][0]
    o = O.Oracle(prm)
    n = prm["n"]
    hx, hy, hz = prm["h"]
    kv = np.array([2 * np.pi * 2 / (n * hx), 2 * np.pi * 1 / (n * hy), 2 * np.pi * 3 / (n * hz)])
    z, y, x = np.meshgrid(np.arange(n) * hz, np.arange(n) * hy, np.arange(n) * hx, indexing="ij")
    u = np.cos(kv[0] * x + kv[1] * y + kv[2] * z)
    o.set_field(0, u)
    Nn = o.nonlinear(0)
    s = np.array([np.sin(kv[0] * hx) / hx, np.sin(kv[1] * hy) / hy, np.sin(kv[2] * hz) / hz])
    sig = np.array([1.0, -1.0, -1.0])
    coef = -sum(s[a] * s[b] * sig[a] * sig[b] * lam[a, b] for a in range(3) for b in range(3))
    expect = -prm["epn2"] * prm["LE"] * coef * u
    inner = slice(2, n - 2)
    assert np.abs(Nn[inner, inner, inner] - expect[inner, inner, inner]).max() < 1e-12


# ---------------------------------------------------------------- fixed points
@pytest.mark.parametrize("etd2", [0, 1])
def test_uniform_landau_root_is_fixed_point(etd2):
    """kk = 1, no elasticity: truncated operator + outside coupling = periodic
    Laplacian, so a uniform root of f' stays put (pins the coupling term and
    its missing kk factor)."""
    a2, a3, a4 = OMEGA_LANDAU
    prm = small_params(n=12, nac=2, etd2=etd2, landau=OMEGA_LANDAU, KE=42.0, epn2=0.6955, dt=0.01, h=(2.0, 2.0, 2.0))
    c = (a3 + np.sqrt(a3 * a3 - 4 * a2 * a4)) / (2 * a4)
    o = O.Oracle(prm)
    n = prm["n"]
    o.set_field(0, np.full((n, n, n), c))
    o.set_field(1, np.zeros((n, n, n)))
    o.step(3)
    assert np.abs(o.get_field(0) - c).max() < 1e-12
    assert np.abs(o.get_field(1)).max() < 1e-14


@pytest.mark.parametrize("etd2", [0, 1])
def test_uniform_concentration_is_fixed_point(etd2):
    prm = small_params(n=12, nac=1, nch=1, etd2=etd2, dt=0.005)
    o = O.Oracle(prm)
    n = prm["n"]
    o.set_field(0, np.zeros((n, n, n)))
    o.set_field(1, np.full((n, n, n), 0.1))
    o.step(3)
    assert np.abs(o.get_field(1) - 0.1).max() < 1e-12


def test_kk_breaks_fixed_point():
    """With kk != 1 the reference's coupling (no kk) no longer matches H."""
    a2, a3, a4 = OMEGA_LANDAU
    prm = small_params(n=12, nac=1, landau=OMEGA_LANDAU, KE=42.0, kk=(2.5, 2.5, 2.5))
    c = (a3 + np.sqrt(a3 * a3 - 4 * a2 * a4)) / (2 * a4)
    o = O.Oracle(prm)
    o.set_field(0, np.full((12, 12, 12), c))
    o.step(1)
    assert np.abs(o.get_field(0) - c).max() > 1e-6


# ---------------------------------------------------------------- ETD1 exactness
def test_etd1_linear_step_is_exact_expm():
    """N = outside coupling only (f' = 0, KE = 0): one ETD1 step solves
    du/dt = A u + b exactly, A = LE epn2 (kk-weighted truncated Laplacian),
    b = LE epn2 u_out / h^2 frozen: u1 = e^{dt A} u + A^-1 (e^{dt A} - I) b."""
    n, g = 8, 2
    N = n - 2 * g
    prm = small_params(n=n, nac=1, landau=(0.0, 0.0, 0.0), KE=0.0, kk=(1.3, 0.8, 1.1), epn2=0.7,
                       h=(1.0, 0.9, 1.2), dt=0.05)
    o = O.Oracle(prm)
    rng = np.random.default_rng(11)
    P = rng.standard_normal((N, N, N))
    idx = (np.arange(n) - g) % N
    u0 = P[np.ix_(idx, idx, idx)]
    o.set_field(0, u0)
    o.step(1)
    T = np.diag(-2.0 * np.ones(n)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    I = np.eye(n)
    hx, hy, hz = prm["h"]
    kx, ky, kz = prm["kk"]
    # flat index (k*n + j)*n + i -> kron(Z, Y, X)
    A = kx * np.kron(I, np.kron(I, T)) / hx ** 2 + ky * np.kron(I, np.kron(T, I)) / hy ** 2 + kz * np.kron(T, np.kron(I, I)) / hz ** 2
    A *= prm["LE"] * prm["epn2"]
    b = np.zeros((n, n, n))
    outside = lambda e: (e - g) % N  # interior image index into P
    for k in range(n):
        for j in range(n):
            b[k, j, 0] += P[outside(k), outside(j), outside(-1)] / hx ** 2
            b[k, j, n - 1] += P[outside(k), outside(j), outside(n)] / hx ** 2
            b[k, 0, j] += P[outside(k), outside(-1), outside(j)] / hy ** 2
            b[k, n - 1, j] += P[outside(k), outside(n), outside(j)] / hy ** 2
            b[0, k, j] += P[outside(-1), outside(k), outside(j)] / hz ** 2
            b[n - 1, k, j] += P[outside(n), outside(k), outside(j)] / hz ** 2
    b = prm["LE"] * prm["epn2"] * b.reshape(-1)
    E = expm(prm["dt"] * A)
    u1 = E @ u0.reshape(-1) + np.linalg.solve(A, (E - np.eye(n ** 3)) @ b)
    assert np.abs(o.get_field(0).reshape(-1) - u1).max() < 1e-12


def test_etd2_is_second_order():
    prm = small_params(n=12, nac=2, etd2=1, dt=0.08, landau=(0.2, 3.0, 3.0), KE=3.2)
    f = noise_fields(prm, amp=0.4)

    def run(dt, steps):
        p = dict(prm, dt=dt)
        o = O.Oracle(p)
        for v in range(2):
            o.set_field(v, f[v])
        o.step(steps)
        return o.get_field(0)

    ref = run(0.005, 64)
    e1 = np.abs(run(0.04, 8) - ref).max()
    e2 = np.abs(run(0.02, 16) - ref).max()
    assert 3.0 < e1 / e2 < 5.0


def test_etd1_is_first_order():
    prm = small_params(n=12, nac=2, etd2=0, dt=0.08)
    f = noise_fields(prm, amp=0.4)

    def run(dt, steps):
        o = O.Oracle(dict(prm, dt=dt))
        for v in range(2):
            o.set_field(v, f[v])
        o.step(steps)
        return o.get_field(0)

    ref = run(0.0025, 128)
    e1 = np.abs(run(0.04, 8) - ref).max()
    e2 = np.abs(run(0.02, 16) - ref).max()
    assert 1.6 < e1 / e2 < 2.4


# ---------------------------------------------------------------- elasticity
def test_b_k0_uses_y_direction_and_is_symmetric_even():
    prm = small_params(n=10, nac=3, elastic=1)
    o = O.Oracle(prm)
    N = 6
    C4 = np.zeros((3, 3, 3, 3))
    vmap = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (1, 2): 3, (2, 1): 3, (0, 2): 4, (2, 0): 4, (0, 1): 5, (1, 0): 5}
    Cv = prm["C"]
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    C4[i, j, k, l] = Cv[vmap[(i, j)], vmap[(k, l)]]
    eps = prm["eps0"]

    def B_direct(nv, p, q):
        sig = np.einsum("ijkl,pkl->pij", C4, eps)
        Om = np.linalg.inv(np.einsum("ijkl,j,l->ik", C4, nv, nv))
        return np.einsum("ijkl,ij,kl->", C4, eps[p], eps[q]) - nv @ sig[p] @ Om @ sig[q] @ nv

    for p in range(3):
        for q in range(3):
            assert abs(o.bpq(N, 0, 0, 0, p, q) - B_direct(np.array([0.0, 1.0, 0.0]), p, q)) < 1e-12
            for (gx, gy, gz) in [(1, 2, 3), (5, 0, 1), (3, 3, 3), (2, 5, 4)]:
                b = o.bpq(N, gx, gy, gz, p, q)
                mirror = o.bpq(N, (-gx) % N, (-gy) % N, (-gz) % N, p, q)
                sx = [gx if gx < N // 2 else gx - N, gy if gy < N // 2 else gy - N, gz if gz < N // 2 else gz - N]
                nv = np.array(sx, dtype=float)
                nv /= np.linalg.norm(nv)
                assert abs(b - B_direct(nv, p, q)) < 1e-12
                assert abs(b - o.bpq(N, gx, gy, gz, q, p)) < 1e-12
                if all(t != N // 2 for t in (gx, gy, gz)):
                    assert abs(b - mirror) < 1e-12


def test_elastic_uniform_field_k0_kat():
    """Uniform eta: felas_p = sum_q B_pq(n = y) eta_q on the interior (this fixes
    the backward-transform normalisation to 1/N^3)."""
    prm = small_params(n=10, nac=3, elastic=1)
    o = O.Oracle(prm)
    n, g = 10, 2
    etas = [0.3, -0.2, 0.7]
    for v in range(3):
        o.set_field(v, np.full((n, n, n), etas[v]))
    o.elastic()
    N = n - 2 * g
    inner = slice(g, n - g)
    for p in range(3):
        expect = sum(o.bpq(N, 0, 0, 0, p, q) * etas[q] for q in range(3))
        assert np.abs(o.felas(p)[inner, inner, inner] - expect).max() < 1e-12


def test_elastic_brute_force_convolution():
    prm = small_params(n=8, nac=2, elastic=1)
    o = O.Oracle(prm)
    n, g = 8, 2
    N = n - 2 * g
    f = noise_fields(prm, amp=0.5)
    for v in range(2):
        o.set_field(v, f[v])
    o.elastic()
    idx = np.arange(N)
    W = np.exp(-2j * np.pi * np.outer(idx, idx) / N)
    Wi = np.conj(W)
    spec = [np.einsum("az,by,cx,zyx->abc", W, W, W, fi[g:n - g, g:n - g, g:n - g]) for fi in f]
    for p in range(2):
        acc = np.zeros((N, N, N), dtype=complex)
        for q in range(2):
            B = np.array([[[o.bpq(N, x, y, z, p, q) for x in range(N)] for y in range(N)] for z in range(N)])
            acc += B * spec[q]
        real = np.einsum("az,by,cx,zyx->abc", Wi, Wi, Wi, acc).real / N ** 3
        assert np.abs(o.felas(p)[g:n - g, g:n - g, g:n - g] - real).max() < 1e-12


# ---------------------------------------------------------------- diagnostics
def test_diag_single_cell_by_hand():
    """n = 2g + 2: one energy cell; check against the trilinear-cell formulas."""
    prm = small_params(n=6, nac=2, nch=1, landau=(0.4, 1.5, 2.0), epn2=0.8, h=(0.5, 1.0, 2.0))
    o = O.Oracle(prm)
    rng = np.random.default_rng(5)
    f = [rng.standard_normal((6, 6, 6)) for _ in range(3)]
    for v in range(3):
        o.set_field(v, f[v])
    st, couple = o.diag()
    hx, hy, hz = prm["h"]
    h3 = hx * hy * hz
    cell = [fi[2:4, 2:4, 2:4] for fi in f]  # [k, j, i]
    means = [c.mean() for c in cell]
    grads = []
    for c in cell:
        gx = (c[:, :, 1] - c[:, :, 0]).mean() / hx
        gy = (c[:, 1, :] - c[:, 0, :]).mean() / hy
        gz = (c[1, :, :] - c[0, :, :]).mean() / hz
        grads.append(np.array([gx, gy, gz]))
    a2, a3, a4 = prm["landau"]
    e = 0.0
    for m in range(2):
        e += prm["epn2"] * grads[m] @ prm["lam"][m] @ grads[m] / 2
    s2 = means[0] ** 2 + means[1] ** 2
    s3 = means[0] ** 3 + means[1] ** 3
    e += a2 / 2 * s2 - a3 / 3 * s3 + a4 / 4 * s2 * s2
    assert abs(st[0, 3] - e * h3) < 1e-12 * max(1, abs(e))
    assert st[1, 3] == st[0, 3]
    ech = prm["epn2"] * grads[2] @ grads[2] / 2 + (means[2] ** 2 - 1) ** 2 / 4
    assert abs(st[2, 3] - ech * h3) < 1e-12 * max(1, abs(ech))
    ecp = (ech * h3 + prm["epn2"] * (grads[0] @ grads[0] + grads[2] @ grads[2]) / 2) * h3
    assert abs(couple - ecp) < 1e-12 * max(1, abs(ecp))
    for v in range(3):
        inner = f[v][2:4, 2:4, 2:4]
        assert st[v, 0] == inner.max() and st[v, 1] == inner.min()
        assert abs(st[v, 2] - (inner + 1.0).sum() * h3) < 1e-12


def test_step_pair_equals_full_and_half_steps():
    """orc_step_pair (the adaptive attempt's shared-N pair) is orc_step at dt
    and orc_step at dt/2 from the same state, bit for bit."""
    prm = small_params(n=12, nac=2, nch=1, elastic=1, aniso=1, dt=0.02)
    f = noise_fields(prm, amp=0.3)
    a, b, c = O.Oracle(prm), O.Oracle(dict(prm)), O.Oracle(dict(prm, dt=0.01))
    for v, x in enumerate(f):
        a.set_field(v, x)
        b.set_field(v, x)
        c.set_field(v, x)
    full = a.step_pair(0.02)
    b.step()
    c.step()
    for v in range(3):
        assert np.array_equal(full[v], b.get_field(v).reshape(-1))
        assert np.array_equal(a.get_field(v), c.get_field(v))


def test_mlp_oracle_gradient_by_finite_differences():
    import mlp_oracle as M
    rng = np.random.default_rng(3)
    p = M.init(5)
    p[M.NPAR - 1] = 0.3
    x = rng.standard_normal((7, M.NF))
    t = rng.standard_normal(7)
    loss, g = M.loss_grad(p, x, t)
    idx = list(rng.choice(M.NPAR, 60, replace=False)) + [M.NPAR - 1, 0, 1024, 1024 + 64]
    for i in idx:
        e = np.zeros(M.NPAR)
        e[i] = 1e-6
        fd = (M.loss_grad(p + e, x, t)[0] - M.loss_grad(p - e, x, t)[0]) / 2e-6
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(g[i])), i
    # GELU against its definition at a few points (erf-based, exact)
    a = np.array([-3.0, -0.5, 0.0, 0.7, 2.5])
    from math import erf, sqrt
    assert np.allclose(M.gelu(a), [0.5 * v * (1 + erf(v / sqrt(2))) for v in a], rtol=0, atol=1e-15)


def test_adaptive_restatement_with_everything_accepted_is_half_steps():
    """With a huge tolerance, no extrapolation, no energy test and fixed
    dt the adaptive restatement is plain stepping at dt/2."""
    import adaptive as AD
    dt = 2.0 ** -6  # binary-exact times, so no clipped last step
    prm = small_params(n=12, nac=2, nch=0, elastic=1, dt=dt)
    f = noise_fields(prm, amp=0.3)
    a, b = O.Oracle(prm), O.Oracle(dict(prm, dt=dt / 2))
    for v, x in enumerate(f):
        a.set_field(v, x)
        b.set_field(v, x)
    t, rows, cnt = AD.run(a, t_end=3 * dt, dt0=dt, dt_min=1e-6, dt_max=1.0, tol=1e9, grow=1.0, shrink=1.0,
                          extrapolate=False, energy_test=False)
    assert cnt["accepted"] == 3 and cnt["attempts"] == 3 and t == 3 * dt
    for _ in range(6):
        b.step()
    for v in range(2):
        assert np.abs(a.get_field(v) - b.get_field(v)).max() == 0.0


def test_seeded_nuclei_grow_and_energy_falls():
    """Physics of the coarsening workload on the FP64 oracle at 32^3: nuclei
    above the critical radius grow (ordered volume up), the scheme lowers the
    check_soln_new energy, and no variant decays to the beta state."""
    from pyetd import workloads as W
    prm = W.scaled(W.beta_omega_256_coarsening(), 36, precision=0, dt=0.1)
    f = W.initial_fields(prm)
    o = O.Oracle(prm)
    for v, x in enumerate(f):
        o.set_field(v, x)
    g = prm["g"]
    inner = (slice(g, -g),) * 3
    vol0 = sum(float((x[inner] > 0.5).mean()) for x in f)
    e = []
    for _ in range(4):
        o.step(25)
        sd, _ = o.diag()
        e.append(sd[0, 3])
    vol = sum(float((o.get_field(v)[inner] > 0.5).mean()) for v in range(4))
    assert vol > 2.0 * vol0
    assert all(b < a for a, b in zip(e, e[1:]))
    assert max(float(o.get_field(v).max()) for v in range(4)) > 0.85


def test_adaptive_restatement_refuses_non_finite_states():
    """A NaN state raises before any step; a state whose candidates overflow
    (one cell at 1e75: finite energy, infinite candidate energy) is rejected
    with the factor shrink down to dt_min and then raises, with the fields
    restored."""
    import adaptive as AD
    prm = small_params(n=12, nac=2, nch=1, dt=0.01)
    f = noise_fields(prm, amp=0.05)
    o = O.Oracle(prm)
    g = [x.copy() for x in f]
    g[0][5, 6, 7] = np.nan
    for v, x in enumerate(g):
        o.set_field(v, x)
    with pytest.raises(AD.NonFinite, match="initial state"):
        AD.run(o, t_end=0.1, dt0=0.01, dt_min=1e-4, dt_max=0.1, tol=1e-4)
    g = [x.copy() for x in f]
    g[0][6, 6, 6] = 1e75
    for v, x in enumerate(g):
        o.set_field(v, x)
    with pytest.raises(AD.NonFinite, match="dt_min"):
        AD.run(o, t_end=0.1, dt0=0.01, dt_min=1e-4, dt_max=0.1, tol=1e-4)
    for v in range(3):
        assert np.array_equal(o.get_field(v), g[v])


# ---------------------------------------------------------------- CH pinning
# The reference's CH_f is broken (formula.c:40-56); the oracle uses the
# double well W(c) = (c^2 - 1)^2 / 4, CH_f = W'(c) = c^3 - c.  Three checks
# fixed by the mathematics, each failing for a planted sign flip of CH_f or
# for lambda in place of lambda^2 in the CH operator (oracle mutations).
CH_PRM = dict(LCI=0.7, KCI=1.3, epn2=0.6, h=(1.0, 0.9, 1.2))


def W_prime(c):
    """dW/dc of W = (c^2 - 1)^2 / 4 by a central difference of W itself."""
    W = lambda x: (x * x - 1.0) ** 2 / 4.0
    e = 1e-5
    return (W(c + e) - W(c - e)) / (2 * e)


def ch_bump_error():
    """t(c) = -LCI KCI c + LCI W'(c) pointwise, up to a constant: a uniform
    field c0 plus delta at one interior cell p has out(p + x) = (t(c0 + delta)
    - t(c0)) / hx^2 (the Laplacian of a one-cell bump), away from the box
    edges where the outside-value terms enter."""
    worst = 0.0
    for c0 in (-0.8, 0.1, 0.65):
        for delta in (-0.3, 0.05, 0.4):
            prm = dict(small_params(n=12, nac=1, nch=1), **CH_PRM)
            o = O.Oracle(prm)
            f = np.full((12, 12, 12), c0)
            f[6, 6, 6] += delta
            o.set_field(0, np.zeros((12, 12, 12)))
            o.set_field(1, f)
            out = o.nonlinear(1)
            hx = prm["h"][0]
            t = lambda c: -prm["LCI"] * prm["KCI"] * c + prm["LCI"] * W_prime(c)
            got = out[6, 6, 7] * hx * hx
            worst = max(worst, abs(got - (t(c0 + delta) - t(c0))) / max(1e-3, abs(t(c0 + delta) - t(c0))))
    return worst


def ch_fourier_error():
    """c = c0 + A cos(k.x): W'(c) = c^3 - c has cos(m k.x) components
    m = 1: 3 c0^2 A + 3 A^3 / 4 - A, m = 2: 3 c0 A^2 / 2, m = 3: A^3 / 4; the
    7-point Laplacian multiplies cos(m k.x) by -sum_a 4 sin^2(m k_a h_a / 2) / h_a^2.
    Interior cells two away from the box edges."""
    n = 16
    prm = dict(small_params(n=n, nac=1, nch=1), **CH_PRM)
    hx, hy, hz = prm["h"]
    kv = np.array([2 * np.pi * 2 / (n * hx), 2 * np.pi * 1 / (n * hy), 2 * np.pi * 3 / (n * hz)])
    z, y, x = np.meshgrid(np.arange(n) * hz, np.arange(n) * hy, np.arange(n) * hx, indexing="ij")
    th = kv[0] * x + kv[1] * y + kv[2] * z
    c0, A = 0.15, 0.4
    o = O.Oracle(prm)
    o.set_field(0, np.zeros((n, n, n)))
    o.set_field(1, c0 + A * np.cos(th))
    out = o.nonlinear(1)
    lam = lambda m: -sum(4.0 * np.sin(m * kv[a] * h / 2) ** 2 / h ** 2 for a, h in enumerate(prm["h"]))
    coef = {1: 3 * c0 ** 2 * A + 0.75 * A ** 3 - A, 2: 1.5 * c0 * A ** 2, 3: 0.25 * A ** 3}
    L, K = prm["LCI"], prm["KCI"]
    expect = L * (lam(1) * (coef[1] - K * A) * np.cos(th) + lam(2) * coef[2] * np.cos(2 * th) +
                  lam(3) * coef[3] * np.cos(3 * th))
    inner = slice(2, n - 2)
    return np.abs(out[inner, inner, inner] - expect[inner, inner, inner]).max() / np.abs(expect).max()


def ch_expm_error():
    """One ETD1 step of the CH field is exact for du/dt = L u + N(u0) with N
    frozen at u0: u1 = e^{dt L} u0 + L^-1 (e^{dt L} - I) N(u0), where
    L = -LCI (epn2 Lt^2 - KCI Lt), Lt the truncated (Dirichlet) Kronecker
    Laplacian of the extended box, and N(u0) = R(u0) - L u0 with R the
    periodic CH right-hand side LCI Lap(W'(c) - epn2 Lap c) at the periodic
    image of every extended cell (truncated operator + outside terms = the
    periodic operator)."""
    n, g = 8, 2
    N = n - 2 * g
    prm = dict(small_params(n=n, nac=1, nch=1, dt=0.01), **CH_PRM)
    hx, hy, hz = prm["h"]
    rng = np.random.default_rng(21)
    P = 0.2 + 0.3 * rng.standard_normal((N, N, N))  # [z][y][x]
    idx = (np.arange(n) - g) % N
    u0 = P[np.ix_(idx, idx, idx)]
    o = O.Oracle(prm)
    o.set_field(0, np.zeros((n, n, n)))
    o.set_field(1, u0)
    o.step(1)
    T = np.diag(-2.0 * np.ones(n)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    I = np.eye(n)
    Lt = np.kron(I, np.kron(I, T)) / hx ** 2 + np.kron(I, np.kron(T, I)) / hy ** 2 + np.kron(T, np.kron(I, I)) / hz ** 2
    Lop = -prm["LCI"] * (prm["epn2"] * Lt @ Lt - prm["KCI"] * Lt)

    def lap(a):
        return ((np.roll(a, 1, 2) + np.roll(a, -1, 2) - 2 * a) / hx ** 2 + (np.roll(a, 1, 1) + np.roll(a, -1, 1) - 2 * a) / hy ** 2
                + (np.roll(a, 1, 0) + np.roll(a, -1, 0) - 2 * a) / hz ** 2)

    R = prm["LCI"] * lap(W_prime(P) - prm["epn2"] * lap(P))
    Nu = R[np.ix_(idx, idx, idx)].reshape(-1) - Lop @ u0.reshape(-1)
    E = expm(prm["dt"] * Lop)
    u1 = E @ u0.reshape(-1) + np.linalg.solve(Lop, (E - np.eye(n ** 3)) @ Nu)
    return np.abs(o.get_field(1).reshape(-1) - u1).max() / np.abs(u1).max()


@pytest.mark.parametrize("check", [ch_bump_error, ch_fourier_error, ch_expm_error])
def test_ch_oracle_pinned(check):
    O.set_mutation(0)
    err = check()
    # W' by a central difference carries ~1e-11; dense expm ~1e-13
    assert err < 1e-8, err


@pytest.mark.parametrize("mutation,checks", [(1, [ch_bump_error, ch_fourier_error, ch_expm_error]),
                                             (2, [ch_expm_error])])
def test_ch_checks_fail_for_planted_mistakes(mutation, checks):
    """Each check catches the mistakes it is meant to catch: a sign-flipped
    CH_f (all three) and lambda in place of lambda^2 in the CH linear
    operator (the expm step; the nonlinear term does not contain it)."""
    O.set_mutation(mutation)
    try:
        errs = [chk() for chk in checks]
    finally:
        O.set_mutation(0)
    assert all(e > 1e-3 for e in errs), errs


# ---------------------------------------------------------------- beta -> alpha output formats
def test_field2b_known_answers():
    import formats as F
    u = np.array([-0.1, 0.0, 0.9, 2.0, 2.5, -1.0, np.nan, 0.25])
    got = F.field2b(u)
    # (u + 0.1) * 30000 truncated: 0, 3000, 30000, 63000; 2.5 -> 78000 wraps to
    # 78000 - 65536; -1 and NaN saturate to 0; 0.25 -> 10500 (0.35 * 30000 is
    # 10499.999..., truncated)
    assert list(got) == [0, 3000, 30000, 63000, 78000 - 65536, 0, 0, int((0.25 + 0.1) * 30000.0)]
    assert got.dtype == np.uint16


def test_downsample_and_detector_pinned():
    import formats as F
    rng = np.random.default_rng(8)
    n, g = 20, 2
    u = rng.uniform(-0.2, 0.4, size=(n, n, n))
    m = F.downsample(u, g)
    inner = u[g:-g, g:-g, g:-g]
    ref = inner.reshape(8, 2, 8, 2, 8, 2).mean(axis=(1, 3, 5))
    assert np.abs(m - ref).max() < 1e-15
    bits, binary = F.detector(u, g)
    assert np.abs(ref.reshape(-1) - 0.1).min() > 1e-12  # no ties between the two sum orders
    assert np.array_equal(binary, (ref.reshape(-1) > 0.1).astype(np.float32))
    assert np.array_equal(bits, np.packbits(binary.astype(bool), bitorder="little"))
    # single cells: the first cell of a byte is bit 0, the eighth bit 7; the
    # threshold is strict
    u = np.zeros((n, n, n))
    u[g:g + 2, g:g + 2, g:g + 2] = 0.3              # cell 0
    u[g:g + 2, g:g + 2, g + 14:g + 16] = 0.3        # cell 7 (x = 7)
    u[g, g + 2, g] = 0.8                            # cell 8: sum 0.8, mean 0.8 / 8 == 0.1 exactly -> off
    u[g, g + 2, g + 2] = np.nextafter(0.8, 1.0)     # cell 9: one ulp above -> on
    bits, binary = F.detector(u, g)
    assert bits[0] == 0b10000001 and bits[1] == 0b00000010 and binary[8] == 0.0 and binary[9] == 1.0

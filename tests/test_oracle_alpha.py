"""Pins the oracle's beta -> alpha physics and variable-mobility CH against
what the reference's own files and the mathematics fix:

- CALPHAD: the lattice stabilities and interaction parameters against the
  values the reference writes beside them (source_beta_to_alpha/src/vars.c:
  146-190 comments), and G_bcc, G_hcp against a general Redlich-Kister
  evaluation of the same parameter tables;
- the driving force (anisotropic_hip.c:2075-2140) as the derivative of an
  independently written free energy (finite differences);
- the grain mask term (anisotropic_hip.c:2206-2210);
- the compact Lambda_ij d_i d_j stencil (anisotropic_hip.c:528-647): exact on
  quadratic fields, and its outside neighbours as the periodic image;
- the variable-mobility divergence (mobility.c:435-722) on closed forms, and
  the assembled CH term (mobility.c:838-922) as C times the constant-mobility
  term where the mobility is uniform."""
import numpy as np
import pytest

import oracle as O
from conftest import small_params


# ---------------------------------------------------------------- CALPHAD
def test_calphad_terms_match_reference_notes():
    # GHSERAL, GHSERTI, GHSERV, GBCCAL, GBCCTI, GHCPAL, GHCPV, BL13_0, BL12_0, BL12_1, BL32_0, HL13_0, HL32_0
    noted = [8.9664e4, 7.2379e3, 2.7046e4, 9.4487e4, 7.4701e3, 9.3179e4, 3.3669e4,
             -85873, -63024, -1965, 8860.5, -91123, 19320]
    got = O.calphad_terms()
    for g, r in zip(got, noted):
        assert abs(g - r) <= 5e-5 * abs(r) + 0.5, (g, r)


def redlich_kister(x, G0, pairs, ternary, RT):
    """G = sum x_i G_i + RT sum x_i ln x_i + sum_{(i,j)} x_i x_j sum_k L_k (x_i - x_j)^k
    + x_1 x_2 x_3 sum_i x_i L^(i) -- the general form, from its parameter tables."""
    G = sum(x[i] * G0[i] for i in x) + RT * sum(x[i] * np.log(x[i]) for i in x)
    for (i, j), L in pairs.items():
        G += x[i] * x[j] * sum(Lk * (x[i] - x[j]) ** k for k, Lk in enumerate(L))
    G += x["Al"] * x["V"] * x["Ti"] * sum(x[e] * L for e, L in ternary.items())
    return G


def test_calphad_gibbs_energies_follow_redlich_kister():
    T, R = 1093.0, 8.31451
    t = O.calphad_terms()
    ser_al, ser_ti, ser_v, bcc_al, bcc_ti, hcp_al, hcp_v = t[:7]
    bcc = redlich_kister({"Al": 0.102, "V": 0.036, "Ti": 1 - 0.102 - 0.036},
                         {"Al": bcc_al, "V": ser_v, "Ti": bcc_ti},
                         {("Al", "V"): [-98000 + 32 * T, 3500 - 5 * T],
                          ("Ti", "V"): [10500 - 1.5 * T, 2000, 1000],
                          ("Al", "Ti"): [-128500 + 39 * T, 6000, 21200]},
                         {"Al": -8000.0, "V": -8000.0, "Ti": -8000.0}, R * T)
    hcp = redlich_kister({"Al": 0.1057, "V": 0.0223, "Ti": 1 - 0.1057 - 0.0223},
                         {"Al": hcp_al, "V": hcp_v, "Ti": ser_ti},
                         {("Al", "V"): [-98000 + 32 * T, 3500 - 5 * T],
                          ("Ti", "V"): [30250 - 10 * T],
                          ("Al", "Ti"): [-133750 + 39 * T, 250, 17250]},
                         {"Al": -60000.0, "Ti": -30000.0, "V": 60000.0}, R * T)
    gb, gh = O.calphad()
    assert abs(gb - bcc) < 1e-9 * abs(bcc) and abs(gh - hcp) < 1e-9 * abs(hcp)
    assert gh < gb  # hcp favoured: the alpha phase grows


def calphad_case(nac=4, n=8, mask=None, aniso=0, lam=None):
    prm = dict(small_params(n=n, nac=nac), free_energy=1, KE=3.0, LE=2.0, epn2=0.048)
    prm["aniso"] = aniso
    if lam is not None:
        prm["lam"] = lam
    if mask is not None:
        prm["mask"] = mask
    return prm


def test_calphad_driving_force_is_derivative_of_free_energy():
    """N_p = LE KE u_p - LE dF/du_p at cells off the box faces, with
    F = (dG P(u_p) + wmega W(u_p)) / gnormal + 2 ALPHA sum_{m<m'} u_m^2 u_m'^2,
    P = u^3 (10 - 15u + 6u^2), W = 4u^2 - 4u^3 + u^4."""
    nac, n = 4, 8
    prm = calphad_case(nac, n)
    rng = np.random.default_rng(3)
    fields = [rng.uniform(-0.2, 1.2, size=(n, n, n)) for _ in range(nac)]
    o = O.Oracle(prm)
    for v, f in enumerate(fields):
        o.set_field(v, f)
    gb, gh = O.calphad()

    def F(u):
        P = u ** 3 * (10 - 15 * u + 6 * u ** 2)
        W = 4 * u ** 2 - 4 * u ** 3 + u ** 4
        s = sum((gh - gb) * P[m] + 100.0 * W[m] for m in range(nac)) / 50000.0
        for m in range(nac):
            for q in range(m + 1, nac):
                s += 2 * 0.02 * u[m] ** 2 * u[q] ** 2
        return s

    cell = (3, 4, 2)
    u = np.array([f[cell] for f in fields])
    for p in range(nac):
        N = o.nonlinear(p)[cell]
        e = np.zeros(nac)
        e[p] = 1e-5
        dF = (F(u + e) - F(u - e)) / 2e-5
        want = prm["LE"] * prm["KE"] * u[p] - prm["LE"] * dF
        assert abs(N - want) < 1e-7 * max(1.0, abs(want)), (p, N, want)


def test_grain_mask_subtracts_half_the_field():
    nac, n = 3, 8
    rng = np.random.default_rng(4)
    fields = [rng.uniform(0, 1, size=(n, n, n)) for _ in range(nac)]
    mask = (rng.uniform(size=(n, n, n)) > 0.4).astype(np.int32)
    outs = []
    for m in (None, mask):
        o = O.Oracle(calphad_case(nac, n, mask=m))
        for v, f in enumerate(fields):
            o.set_field(v, f)
        outs.append([o.nonlinear(p) for p in range(nac)])
    for p in range(nac):
        d = outs[1][p] - outs[0][p]
        assert np.array_equal(d[mask == 1], np.zeros(int(mask.sum())))
        assert np.abs(d[mask == 0] + 0.5 * fields[p][mask == 0]).max() < 1e-13


# ---------------------------------------------------------------- compact stencil
def test_compact_stencil_is_exact_on_quadratics():
    """f2 = sum Lambda_ab d_a d_b u = sum Lambda_ab (A + A^T)_ab for
    u = x^T A x + b.x on cells whose 19-point stencil stays in the box;
    read off N(aniso = 2) - N(aniso = 0) = -epn2 LE f2."""
    n, nac = 8, 2
    rng = np.random.default_rng(5)
    A = rng.uniform(-1, 1, size=(3, 3))  # not symmetric
    b = rng.uniform(-1, 1, size=3)
    lam = rng.uniform(-1, 1, size=(nac, 3, 3))
    h = (0.9, 1.1, 1.3)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    X = np.stack([i * h[0], j * h[1], k * h[2]])  # x, y, z
    u = np.einsum("a...,ab,b...->...", X, A, X) + np.einsum("a,a...->...", b, X)
    res = []
    for aniso in (0, 2):
        prm = dict(small_params(n=n, nac=nac, h=h), aniso=aniso, lam=lam, LE=1.0, epn2=1.0)
        o = O.Oracle(prm)
        o.set_field(0, u)
        o.set_field(1, 0.3 * u)
        res.append(o.nonlinear(0))
    f2 = -(res[1] - res[0])
    want = float(np.sum(lam[0] * (A + A.T)))
    inner = f2[1:-1, 1:-1, 1:-1]
    assert np.abs(inner - want).max() < 1e-9 * max(1.0, abs(want))


def test_compact_stencil_reads_periodic_images_outside_the_box():
    """On a periodic field (ghosts = images) the stencil at a face cell equals
    the stencil of the wrapped interior field, computed with numpy rolls."""
    n, g, nac = 10, 2, 1
    N = n - 2 * g
    rng = np.random.default_rng(6)
    inner = rng.uniform(-1, 1, size=(N, N, N))
    idx = (np.arange(n) - g) % N
    u = inner[np.ix_(idx, idx, idx)]
    lam = rng.uniform(-1, 1, size=(1, 3, 3))
    res = []
    for aniso in (0, 2):
        o = O.Oracle(dict(small_params(n=n, nac=nac), aniso=aniso, lam=lam, LE=1.0, epn2=1.0))
        o.set_field(0, u)
        res.append(o.nonlinear(0))
    f2 = -(res[1] - res[0])
    r = lambda a, dz, dy, dx: np.roll(a, (-dz, -dy, -dx), axis=(0, 1, 2))
    L = lam[0]
    fxx = r(inner, 0, 0, 1) - 2 * inner + r(inner, 0, 0, -1)
    fyy = r(inner, 0, 1, 0) - 2 * inner + r(inner, 0, -1, 0)
    fzz = r(inner, 1, 0, 0) - 2 * inner + r(inner, -1, 0, 0)
    fxy = ((r(inner, 0, 1, 1) - r(inner, 0, 1, -1)) - (r(inner, 0, -1, 1) - r(inner, 0, -1, -1))) / 4
    fxz = ((r(inner, 1, 0, 1) - r(inner, 1, 0, -1)) - (r(inner, -1, 0, 1) - r(inner, -1, 0, -1))) / 4
    fyz = ((r(inner, 1, 1, 0) - r(inner, 1, -1, 0)) - (r(inner, -1, 1, 0) - r(inner, -1, -1, 0))) / 4
    want = (L[0, 0] * fxx + (L[0, 1] + L[1, 0]) * fxy + (L[0, 2] + L[2, 0]) * fxz + L[1, 1] * fyy
            + (L[1, 2] + L[2, 1]) * fyz + L[2, 2] * fzz)
    got = f2[np.ix_(np.arange(n), np.arange(n), np.arange(n))]
    full = want[np.ix_(idx, idx, idx)]  # the extended box's cells, ghosts included
    assert np.abs(got - full).max() < 1e-12


# ---------------------------------------------------------------- variable mobility
def test_div_m_grad_closed_forms():
    n = 9
    h = (0.8, 1.0, 1.25)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    x, y, z = i * h[0], j * h[1], k * h[2]
    inner = (slice(1, -1),) * 3
    # linear a, linear M: div(M grad a) = grad M . grad a
    a = x + 2 * y - z
    M = 1 + 0.1 * x + 0.2 * y + 0.3 * z
    out = O.div_m_grad(M, a, h)
    assert np.abs(out[inner] - (0.1 + 0.4 - 0.3)).max() < 1e-12
    # constant M, quadratic a: M Lap a
    a = x * x + 3 * y * y - 2 * z * z + x * y
    out = O.div_m_grad(np.full((n, n, n), 0.7), a, h)
    assert np.abs(out[inner] - 0.7 * (2 + 6 - 4)).max() < 1e-10


def vm_params(n=10, varmob=1, KCI=0.0):
    prm = small_params(n=n, nac=1, nch=3)
    prm.update(varmob=varmob, LCI=1.0, KCI=KCI)
    return prm


def test_variable_mobility_is_scaled_constant_mobility_where_mobility_is_uniform():
    """eta = 0 everywhere: M_q = C_q = x_q (1 - x_q) D_qA; with LCI = 1 and
    KCI = 0 the variable-mobility term is C_q times the constant-mobility one.
    C_0 = 0.15 * 0.85 (D_CuA = 1 by the reference's normalisation)."""
    n = 10
    rng = np.random.default_rng(8)
    cs = [0.3 + 0.1 * rng.standard_normal((n, n, n)) for _ in range(3)]
    res = {}
    for vm in (0, 1):
        o = O.Oracle(vm_params(n, vm))
        o.set_field(0, np.zeros((n, n, n)))
        for q in range(3):
            o.set_field(1 + q, cs[q])
        o.transfer()
        res[vm] = [o.nonlinear(1 + q) for q in range(3)]
    RT = 8.314472 * 823.0
    dcu = 4.7e-5 * np.exp(-2.44e5 / RT)
    C = [0.15 * 0.85,
         0.01 * 0.99 * 1.49e-4 * np.exp(-2.63e5 / RT) / dcu,
         0.01 * 0.99 * 1.4e-4 * np.exp(-2.56e5 / RT) / dcu]
    for q in range(3):
        assert np.abs(res[1][q] - C[q] * res[0][q]).max() < 1e-12 * np.abs(res[0][q]).max() * max(1, C[q])


def test_variable_mobility_divergence_term_off_the_faces():
    """With eta varying, N_vm - C N_const = -LCI div((M - C) grad(epn2 Lap c - f0))
    on cells two away from the faces (LCI = 1, KCI = 0, f0 = CH_f(c))."""
    n = 12
    rng = np.random.default_rng(9)
    eta = rng.uniform(0, 1, size=(n, n, n))
    c = 0.3 + 0.05 * rng.standard_normal((n, n, n))
    res = {}
    for vm in (0, 1):
        o = O.Oracle(vm_params(n, vm))
        o.set_field(0, eta)
        for q in range(3):
            o.set_field(1 + q, c)
        o.transfer()
        res[vm] = o.nonlinear(1)  # Cu
        eta_f, c = o.get_field(0), o.get_field(1)  # after the ghost fill
    RT = 8.314472 * 823.0
    dcu = 4.7e-5 * np.exp(-2.44e5 / RT)
    DG = 4.3e-5 * np.exp(-2.80e5 / RT) / dcu
    M = 0.15 * 0.85 * ((1 - eta_f) + DG * eta_f)
    C = M.max()
    lap = np.zeros_like(c)
    lap[1:-1, 1:-1, 1:-1] = (c[1:-1, 1:-1, 2:] + c[1:-1, 1:-1, :-2] + c[1:-1, 2:, 1:-1] + c[1:-1, :-2, 1:-1]
                             + c[2:, 1:-1, 1:-1] + c[:-2, 1:-1, 1:-1] - 6 * c[1:-1, 1:-1, 1:-1])
    ft = 1.0 * lap - c * (c * c - 1)
    div = O.div_m_grad(M - C, ft)
    inner = (slice(2, -2),) * 3
    assert np.abs((res[1] - C * res[0])[inner] + div[inner]).max() < 1e-10 * np.abs(div[inner]).max()

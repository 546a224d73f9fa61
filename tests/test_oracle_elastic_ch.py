"""Pins the oracle's elastic force on the CH fields (ELASTIC = 2,
CH_FUNCTION: source_beta_to_omega/src/elastic.c:726-760, calc.c:90-101,
ScLETD.h:153-156) and ELASTIC = 3 (BOTH_FUNCTION, which no branch of
elastic_calculate matches):

- the force on CH field q is the force the AC path computes for the same
  data and eigenstrains (the force itself is pinned by the k = 0 known answer
  and the brute-force convolution in tests/test_oracle.py);
- it enters the CH term as LCI felas inside the Laplacian and minus LCI felas
  at the outside cells: with the force periodic, the difference it makes is
  the periodic Laplacian of LCI felas on every cell;
- ELASTIC = 3 gives the force-free result."""
import numpy as np

import oracle as O
from conftest import small_params


def ch_case(elastic, nac=1, nch=2, n=12, seed=21):
    prm = small_params(n=n, nac=max(nac, nch), elastic=1, seed=seed)  # eigenstrains for max(nac, nch) variants
    prm.update(nac=nac, nch=nch, elastic=elastic)
    return prm


def test_ch_force_equals_ac_force_on_the_same_data():
    n = 12
    rng = np.random.default_rng(22)
    data = [rng.uniform(-0.3, 0.3, size=(n, n, n)) for _ in range(2)]
    a = O.Oracle(ch_case(1, nac=2, nch=0))
    for v in range(2):
        a.set_field(v, data[v])
    a.elastic()
    b = O.Oracle(ch_case(2, nac=1, nch=2))
    b.set_field(0, np.zeros((n, n, n)))
    for q in range(2):
        b.set_field(1 + q, data[q])
    b.elastic()
    for q in range(2):
        assert np.array_equal(a.felas(q), b.felas(q))


def test_ch_force_enters_as_periodic_laplacian_of_lci_felas():
    """One CH field, so the shared force buffer's ghosts (each fill reads the
    previous one's edges and corners) hold this force's own images after a
    third evaluation."""
    n, g = 12, 2
    N = n - 2 * g
    rng = np.random.default_rng(23)
    c = 0.2 + 0.1 * rng.standard_normal((n, n, n))
    res = {}
    for el in (0, 2):
        prm = ch_case(el, nac=1, nch=1)
        prm["LCI"] = 0.7
        o = O.Oracle(prm)
        o.set_field(0, np.zeros((n, n, n)))
        o.set_field(1, c)
        o.transfer()
        if el:
            for _ in range(3):
                o.elastic()
            fe = o.felas(0)
        res[el] = o.nonlinear(1)
    idx = (np.arange(n) - g) % N
    assert np.array_equal(fe, fe[g:n - g, g:n - g, g:n - g][np.ix_(idx, idx, idx)])  # ghosts are images
    h = small_params(n=n, nac=1)["h"]
    inner = 0.7 * fe[g:n - g, g:n - g, g:n - g]
    lap = sum((np.roll(inner, 1, ax) + np.roll(inner, -1, ax) - 2 * inner) / h[2 - ax] ** 2 for ax in range(3))
    want = lap[np.ix_(idx, idx, idx)]
    assert np.abs((res[2] - res[0]) - want).max() < 1e-12 * np.abs(want).max()


def test_elastic_both_computes_no_force():
    n = 12
    rng = np.random.default_rng(24)
    f = [rng.uniform(-0.3, 0.3, size=(n, n, n)) for _ in range(3)]
    out = {}
    for el in (0, 3):
        prm = ch_case(el, nac=2, nch=1)
        o = O.Oracle(prm)
        for v in range(3):
            o.set_field(v, f[v])
        o.step(2)
        out[el] = [o.get_field(v) for v in range(3)]
    for v in range(3):
        assert np.array_equal(out[0][v], out[3][v])

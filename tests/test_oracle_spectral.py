"""The oracle's spectral helpers (ch_opt.f90:28-152: mk_cdf, search_1D,
search_2D, lin_inter_1D, lin_inter_2D, sample) against the reference's own
ch_opt module compiled from its sources: bit-exact on the committed golden
vectors (tests/golden/reference_vectors.json) and, where oracle/_ref is
built, on random spectra; plus closed forms that fix the index conventions:
a uniform spectrum samples x1 + u (xK - x1) for every u (the extrapolation
below the first CDF node included), and a linear density, whose trapezoid
CDF is exact, maps u = CDF(node) onto the node."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_ffi as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))["spectral"]


def hexs(a):
    return np.array([float.fromhex(x) for x in a])


def _table():
    return O.spectrum_table(GOLD["x"], GOLD["y"])


def test_mk_cdf_matches_reference_vectors():
    cdf = O.mk_cdf(_table(), cdf_last=GOLD["cdf_last_in"])
    assert np.array_equal(cdf, hexs(GOLD["cdf"]))


def test_sample_matches_reference_vectors():
    tab = _table()
    cdf = hexs(GOLD["cdf"])
    assert np.array_equal(O.sample_seq(tab, cdf, GOLD["sample_seed"], 200), hexs(GOLD["samples"]))


def test_search_and_interpolation_match_reference_vectors():
    L = O.lib()
    tab = _table()
    cdf = hexs(GOLD["cdf"])
    n = len(GOLD["x"])
    for i, v in enumerate(GOLD["values"]):
        k1 = L.or_search_1d(n, cdf, v)
        k2 = L.or_search_2d(n, tab, v)
        assert (k1, k2) == (GOLD["search_1d"][i], GOLD["search_2d"][i]), v
        c1, c2 = min(max(k1, 1), n - 2), min(max(k2, 1), n - 1)
        assert L.or_lin_inter_1d(tab, cdf, v, n, c1) == float.fromhex(GOLD["lin_inter_1d"][i]), v
        assert L.or_lin_inter_2d(tab, v, n, c2) == float.fromhex(GOLD["lin_inter_2d"][i]), v


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("n", [3, 4, 17, 64])
def test_random_spectra_match_live_reference(n):
    rng = np.random.default_rng(n)
    x = np.sort(rng.uniform(300.0, 1100.0, n))
    y = rng.uniform(0.0, 2.0, n)
    tab = O.spectrum_table(x, y)
    R = O.ref()
    cr = np.zeros(n)
    R.ref_mk_cdf(n, tab, cr)
    assert np.array_equal(O.mk_cdf(tab), cr)
    s = C.c_int(-31 - n)
    wr = np.zeros(500)
    R.ref_sample(n, 500, tab, cr, C.byref(s), wr)
    assert np.array_equal(O.sample_seq(tab, cr, -31 - n, 500), wr)
    vals = np.ascontiguousarray(np.concatenate([rng.uniform(0, 1, 50), rng.uniform(280, 1120, 50), x[:5]]))
    nl = np.zeros(2 * len(vals), np.int32)
    R.ref_search(n, tab, cr, len(vals), vals, nl)
    L = O.lib()
    assert [L.or_search_1d(n, cr, v) for v in vals] == nl[:len(vals)].tolist()
    assert [L.or_search_2d(n, tab, v) for v in vals] == nl[len(vals):].tolist()


def test_uniform_spectrum_samples_linearly():
    """y = 1 on K rows: cdf(j) = (x_{j+1} - x_1) / (x_K - x_1), so the
    inversion is x_1 + u (x_K - x_1) on every bracket, and below cdf(1) the
    extrapolation along the second row's bracket is the same line."""
    L = O.lib()
    x = np.array([500.0, 520.0, 560.0, 580.0, 660.0, 700.0])
    tab = O.spectrum_table(x, np.ones_like(x))
    cdf = O.mk_cdf(tab)
    np.testing.assert_allclose(cdf[:-1], (x[1:] - x[0]) / (x[-1] - x[0]), rtol=1e-15)
    for u in np.linspace(1e-6, 1 - 1e-6, 101):
        assert abs(L.or_sample_u(tab, len(x), cdf, u) - (x[0] + u * (x[-1] - x[0]))) < 1e-10


def test_linear_density_samples_its_nodes():
    """Density y = x on [0, 1] (trapezoid exact): CDF(x_j) = x_j^2, and the
    sample at u = x_j^2 is x_j for the nodes the bracket convention reaches."""
    L = O.lib()
    x = np.linspace(0.0, 1.0, 11)
    tab = O.spectrum_table(x, x)
    cdf = O.mk_cdf(tab)
    np.testing.assert_allclose(cdf[:-1], x[1:] ** 2, rtol=1e-14)
    for j in range(2, 10):
        assert abs(L.or_sample_u(tab, len(x), cdf, x[j] ** 2) - x[j]) < 1e-12, j
    # a plausible off-by-one (bracket rows nlow, nlow+1) would give x_{j-1}
    assert abs(L.or_sample_u(tab, len(x), cdf, x[5] ** 2) - x[4]) > 0.05

"""The library's spectral helpers (tam_spectrum_*, ch_opt.f90:28-152).

CPU: tam_spectrum_cdf (host arithmetic in libtam.so) equals the oracle's
mk_cdf bit for bit on the golden spectrum and on random ones, and the entry
points reject bad tables.  GPU (-m gpu): tam_spectrum_sample equals the
oracle's sample() fed the same Philox uniforms, bit for bit (FP64, the
reference's operation order), and its draws follow the CDF the inversion
defines (Kolmogorov-Smirnov); tam_spectrum_interp equals search_2D +
lin_inter_2D."""
import json
import math
import os

import numpy as np
import pytest

import oracle_ffi as O
from tam_amd import lib, spectral

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))["spectral"]


def _hex(a):
    return np.array([float.fromhex(x) for x in a])


def test_library_cdf_equals_the_oracle_and_reference_vectors():
    tab = np.array([GOLD["x"], GOLD["y"]])
    cdf = spectral.mk_cdf(tab, cdf_last=GOLD["cdf_last_in"])
    assert np.array_equal(cdf, _hex(GOLD["cdf"]))
    rng = np.random.default_rng(4)
    for n in (2, 3, 9, 100):
        x = np.sort(rng.uniform(300, 1100, n))
        y = rng.uniform(0.01, 2.0, n)
        assert np.array_equal(spectral.mk_cdf(np.array([x, y])), O.mk_cdf(O.spectrum_table(x, y)))


def test_bad_tables_are_rejected():
    with pytest.raises(lib.TamError, match="ascend"):
        spectral.mk_cdf(np.array([[1.0, 1.0, 2.0], [1.0, 1.0, 1.0]]))
    with pytest.raises(lib.TamError, match="too short"):
        spectral.mk_cdf(np.array([[1.0], [1.0]]))
    with pytest.raises(lib.TamError, match="positive area"):
        spectral.mk_cdf(np.array([[1.0, 2.0], [0.0, 0.0]]))
    with pytest.raises(lib.TamError, match="too short"):       # sampling needs three rows
        spectral.sample(np.array([[1.0, 2.0], [1.0, 1.0]]), [1.0, 0.0], 4, seed=1)


def _philox_uniforms(n, seed, stream):
    return np.array([O.lib().or_u01(O.philox((i & 0xffffffff, i >> 32, 0, stream),
                                             (seed & 0xffffffff, seed >> 32))[0]) for i in range(n)])


@pytest.mark.gpu
def test_gpu_sample_equals_oracle_on_the_same_uniforms():
    x, y = np.array(GOLD["x"]), np.array(GOLD["y"])
    tab = np.array([x, y])
    cdf = spectral.mk_cdf(tab)
    n, seed, stream = 4096, 0x1234_5678_9abc, 7
    print("case: spectrum sample, 4096 draws, 12-row spectrum", flush=True)
    w = spectral.sample(tab, cdf, n, seed=seed, stream=stream)
    u = _philox_uniforms(n, seed, stream)
    t = O.spectrum_table(x, y)
    ref = np.array([O.lib().or_sample_u(t, len(x), cdf, v) for v in u])
    assert np.array_equal(w, ref)


@pytest.mark.gpu
def test_gpu_samples_follow_the_inverted_cdf():
    """Above cdf(1) the inversion is the linear interpolant of the CDF nodes
    (cdf(j) at row j+1), so the draws' CDF there is that interpolant: KS
    distance within the 1e-3 significance bound for 2e6 draws."""
    x = np.array([400.0, 430.0, 455.0, 500.0, 520.0, 575.0, 600.0, 650.0, 720.0, 800.0, 910.0, 1000.0])
    y = np.array([0.1, 0.35, 0.8, 1.0, 0.95, 0.6, 0.7, 1.2, 0.9, 0.4, 0.2, 0.05])
    tab = np.array([x, y])
    cdf = spectral.mk_cdf(tab)
    n = 2_000_000
    w = np.sort(spectral.sample(tab, cdf, n, seed=99))
    sel = w >= x[1]
    F = np.interp(w[sel], x[1:], cdf[:-1])                  # the interpolant's CDF at each draw
    emp = (np.nonzero(sel)[0] + 1) / n
    d = np.abs(emp - F).max()
    print("KS distance", d, flush=True)
    assert d < 1.95 / math.sqrt(n)
    assert (w >= x[0] - 50).all() and (w <= x[-1]).all()


@pytest.mark.gpu
def test_gpu_interp_equals_oracle():
    rng = np.random.default_rng(2)
    x = np.sort(rng.uniform(300, 1100, 40))
    y = rng.uniform(0, 5, 40)
    q = np.concatenate([rng.uniform(250, 1150, 5000), x])
    got = spectral.lin_inter_2d(np.array([x, y]), q)
    t = O.spectrum_table(x, y)
    L = O.lib()
    ref = np.array([L.or_lin_inter_2d(t, v, 40, L.or_search_2d(40, t, v)) for v in q])
    assert np.array_equal(got, ref)
    # at the wavelength rows the table's own values come back
    np.testing.assert_allclose(got[5000:-1], y[:-1], rtol=0, atol=1e-12)

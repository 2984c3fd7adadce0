"""Batched instances (tam_batch_*) against the same instances run one at a
time: the transport of every instance in one launch gives each instance the
integer fates of its own tam_mc_run on the same stream, and the same tally up
to the rounding of the deposits that either run sums in an LDS box (16-bit
fixed point, 2^-13 (dx + dy + dz) per unit, at most half a unit per deposit;
the batched kernel's per-wave boxes and the single run's CU-wide box cover
different voxels): bound 1e-4 relative plus 16 units; batched heat steps are
bit-identical with the single-instance step of the same launch shape."""
import numpy as np
import pytest

from tam_amd import configs, lib

pytestmark = pytest.mark.gpu

MANIFEST = [dict(edge=32, mua=0.5, mus=100.0, g=0.9, packets=120_000),
            dict(edge=48, mua=5.0, mus=60.0, g=0.75, packets=40_000),
            dict(edge=64, mua=0.2, mus=250.0, g=0.93, packets=90_000),
            dict(edge=40, mua=2.0, mus=150.0, g=0.8, packets=10_000),
            dict(edge=56, mua=8.0, mus=300.0, g=0.7, packets=70_000)]


def _print(*a):
    print(*a, flush=True)


def _grids():
    probs = configs.batch_problems(MANIFEST, steps=51)
    return probs, [p.grid() for p in probs]


def test_batch_transport_matches_single_runs():
    _print("case: 5 uneven instances, batched transport vs single")
    probs, gs = _grids()
    b = lib.Batch(gs)
    b.set_launch(4, 8, 2, 24)
    npk = [m["packets"] for m in MANIFEST]
    st = b.mc_run(npk, seed=42, stream0=10)
    tb = [g.download(lib.FIELD_TALLY) for g in gs]
    b.close()
    for i, g in enumerate(gs):
        g.set_mc_launch(2, 24)
        s1 = g.mc_run(npk[i], seed=42, stream=10 + i)
        t1 = g.download(lib.FIELD_TALLY)
        for k in ("absorbed", "specular", "missed", "crossings", "interactions", "tally_adds"):
            assert getattr(st[i], k) == getattr(s1, k), (i, k)
        assert list(st[i].escaped) == list(s1.escaped)
        p = probs[i].params
        unit = 2.0 * (p.xmax / p.nx + p.ymax / p.ny + p.zmax / p.nz) / 2 ** 13
        np.testing.assert_allclose(tb[i], t1, rtol=1e-4, atol=16 * unit)
        g.close()


@pytest.mark.parametrize("tile,kz", [(3, 4), (4, 8), (5, 16), (6, 3), (7, 8), (8, 5)])
def test_batch_heat_is_bit_identical(tile, kz):
    _print(f"case: 5 uneven instances, 51 batched steps, tile {tile} kz {kz}")
    probs, ga = _grids()
    gb = [p.grid() for p in probs]
    rng = np.random.default_rng(tile)
    for pa, a, bb in zip(probs, ga, gb):
        n = pa.params.nx
        j = 1e7 * rng.random((n, n, n))
        for g in (a, bb):
            g.upload(lib.FIELD_JMEAN, j)         # also forms the per-step source
    b = lib.Batch(gb)
    b.set_launch(tile, kz, 4, 32)
    b.heat_run(51)
    b.close()
    for a, bb in zip(ga, gb):
        a.set_heat_launch(tile, kz)
        a.heat_run(51)
        Ta, Tb = a.download(lib.FIELD_TEMP), bb.download(lib.FIELD_TEMP)
        assert np.array_equal(Ta, Tb)
        assert Ta.max() > configs.T_BODY + 1e-3
        a.close()
        bb.close()


def test_batch_tuner_keeps_the_state():
    probs, gs = _grids()
    T0 = [g.download(lib.FIELD_TEMP) for g in gs]
    b = lib.Batch(gs)
    shape = b.tune([20_000] * len(gs), heat_reps=2)
    _print("tuned", shape)
    assert 3 <= shape["tile"] <= 8 and shape["kz"] in (4, 8, 16, 32)
    b.close()
    for g, t in zip(gs, T0):
        assert np.array_equal(g.download(lib.FIELD_TEMP), t)
        assert not g.download(lib.FIELD_TALLY).any()
        g.close()

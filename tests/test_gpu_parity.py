"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle.

Tolerances (stated per test):
  * MC, same Philox stream: the FP32 walk follows the FP64 oracle's packet
    histories, so fate fractions agree to 0.1% absolute and per-voxel tallies
    to 1% in L1 (rare FP32/FP64 branch flips re-route single packets).
  * MC, independent streams: per-voxel path-length tallies agree within
    3 sigma of the combined batch-means standard error (>= 99% of voxels with
    enough samples), fate fractions within 0.1%.
  * heat: 1e-5 relative to the FP64 oracle.
  * ablation mask: bit-exact given the identical temperature field.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi as O
from tam_amd import configs, lib

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


# ------------------------------------------------------------------ MC ----

def _two_layer_problem(n=40, packets=200_000):
    p = configs._config_params((n, n, n), (1.0, 1.0, 1.0))
    p.beam = lib.BEAM_TOPHAT
    p.beam_radius = 0.1
    p.n_ext_zhi = 1.0
    p.n_ext_zlo = 1.33
    mat = np.zeros((n, n, n), np.uint8)
    mat[n // 2:, :, :] = 1                       # upper half (k >= n/2) is material 1
    table = [(0.5, 100.0, 0.9, 1.4), (2.0, 80.0, 0.85, 1.4)]
    temp, kappa, density, heatcap = configs._thermal_cube(p, configs.C0_THERMAL)
    return configs.Problem("two-layer", p, mat, table, temp, kappa, density, heatcap, packets=packets)


def _oracle_for(prob, seed, stream=0, beam=None):
    p = prob.params
    n_ext = (0, 0, 0, 0, p.n_ext_zlo, p.n_ext_zhi)
    return O.MCProblem(p.nx, p.ny, p.nz, p.xmax, p.ymax, p.zmax, prob.mat, prob.table,
                       mut=None if prob.rhokap is None else prob.rhokap[1:-1, 1:-1, 1:-1],
                       n_ext=n_ext, beam=p.beam if beam is None else beam, radius=p.beam_radius,
                       rng=0, seed=seed,
                       stream=stream)


def _fractions_oracle(st, n):
    return {"absorbed": st.absorbed / n, "specular": st.specular / n,
            "diffuse_reflected": st.escaped[5] / n, "transmitted": st.escaped[4] / n,
            "lateral": sum(st.escaped[:4]) / n, "missed": st.missed / n}


def test_mc_same_stream_follows_oracle():
    prob = _two_layer_problem()
    _print("case: two-layer 40^3, 2e5 packets, seed 1234, same stream")
    g = prob.grid()
    st = g.mc_run(prob.packets, seed=1234)
    t_gpu = g.download(lib.FIELD_TALLY)
    g.close()
    t_or, ost = _oracle_for(prob, 1234).run(prob.packets)
    fg, fo = st.fractions(), _fractions_oracle(ost, prob.packets)
    _print("gpu", fg, "\noracle", fo)
    for k in fg:
        assert abs(fg[k] - fo[k]) <= 1e-3, k
    assert st.packets == st.absorbed + st.specular + sum(st.escaped) + st.missed + st.killed
    rel = np.abs(t_gpu - t_or).sum() / t_or.sum()
    _print("tally L1 rel diff", rel)
    assert rel < 1e-2
    # the work counters follow the same histories
    assert abs(st.interactions / ost.interactions - 1) < 1e-2


def test_mc_same_stream_through_a_crater():
    """A cylinder of mu_t = 0 voxels below the beam (an ablation crater): the
    packets fly through it freely, neither side tallies there, and the rest
    follows the oracle as in the intact case (same tolerances)."""
    prob = _two_layer_problem(n=40, packets=200_000)
    n = 40
    mua = np.array([t[0] + t[1] for t in prob.table])
    rk = np.zeros((n + 2,) * 3)
    rk[1:-1, 1:-1, 1:-1] = mua[prob.mat]
    zz, yy, xx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    crater = ((xx - n / 2 + 0.5) ** 2 + (yy - n / 2 + 0.5) ** 2 <= 36) & (zz >= n - 12)
    rk[1:-1, 1:-1, 1:-1][crater] = 0.0
    prob.rhokap = rk
    _print("case: two-layer 40^3 with a mu_t = 0 crater, 2e5 packets, seed 4321")
    g = prob.grid()
    st = g.mc_run(prob.packets, seed=4321)
    t_gpu = g.download(lib.FIELD_TALLY)
    g.close()
    t_or, ost = _oracle_for(prob, 4321).run(prob.packets)
    assert not t_gpu[crater].any() and not t_or[crater].any()
    fg, fo = st.fractions(), _fractions_oracle(ost, prob.packets)
    _print("gpu", fg, "\noracle", fo)
    for k in fg:
        assert abs(fg[k] - fo[k]) <= 1e-3, k
    rel = np.abs(t_gpu - t_or).sum() / t_or.sum()
    _print("tally L1 rel diff", rel)
    assert rel < 1e-2
    assert abs(st.crossings / ost.crossings - 1) < 1e-2


def test_mc_launch_shape_does_not_change_results():
    """Packets draw from their own streams whichever lane takes them: every
    (blocks per CU, random-phase batch, LDS tally box off / fixed point / FP64;
    with the box on, 4 and 8 run as 1 and 2 blocks of 1024 lanes) gives the
    same integer fates and the same tally up to FP64 summation order, or, with
    the 16-bit fixed-point box, up to the rounding of each deposit to its unit
    u = 2^-13 (dx + dy + dz): half a unit per deposit at most, unbiased, ~1e-4
    of a voxel-sized path.  Bound used: 1e-4 relative plus 16 u (a voxel that
    took a few short paths has errors of a few u).  The sums are integers, so
    the comparison is deterministic.  The tuner keeps the box's kind and
    leaves the tally clear."""
    prob = _two_layer_problem(n=32, packets=300_000)
    p = prob.params
    unit = 2.0 * (p.xmax / p.nx + p.ymax / p.ny + p.zmax / p.nz) / 2 ** 13
    g = prob.grid()
    res = []
    for bpc, batch, box in ((1, 1, 2), (2, 64, 2), (8, 32, 0), (5, 13, 1), (2, 32, 0), (4, 32, 1), (8, 16, 1),
                            (4, 32, 2)):
        _print(f"case: launch bpc={bpc} batch={batch} box={box} (0 off, 1 fixed point, 2 FP64)")
        g.set_mc_launch(bpc, batch)
        g.set_mc_box(box > 0, fp64=box == 2)
        st = g.mc_run(prob.packets, seed=77, stream=3)
        # 300 000 packets clear the short-launch threshold (a quarter of the
        # launch's lanes) at every shape here, so a box case sums in its box
        assert (st.box_adds > 0) == (box > 0), (bpc, box, st.box_adds)
        res.append((box, st, g.download(lib.FIELD_TALLY)))
    b, t, rate = g.tune_mc(100_000)
    assert 1 <= b <= 64 and 1 <= t <= 64 and rate > 0
    assert not g.download(lib.FIELD_TALLY).any()
    g.close()
    _, st0, t0 = res[0]
    for box, st, tl in res[1:]:
        for k in ("absorbed", "specular", "missed", "crossings", "interactions", "killed"):
            assert getattr(st, k) == getattr(st0, k), k
        assert list(st.escaped) == list(st0.escaped)
        if box == 1:
            np.testing.assert_allclose(tl, t0, rtol=1e-4, atol=16 * unit)
        else:
            np.testing.assert_allclose(tl, t0, rtol=1e-9, atol=1e-12)


def test_short_launch_runs_without_the_box():
    """A launch with fewer packets than a quarter of its lanes (4 blocks per
    CU: one 1024-lane block per CU) runs without the LDS tally box even when
    the box is on, since clearing and flushing it would cost more than it
    saves (the reference's own run: 125 000 packets per cycle).  Its fates and
    tally equal those of the same launch with the box off; a launch of more
    packets at the same shape does use the box."""
    prob = _two_layer_problem(n=32, packets=20_000)
    g = prob.grid()
    g.set_mc_launch(4, 32)
    res = []
    for box in (False, True):
        _print(f"case: 20 000 packets, 4 blocks per CU, box {'on' if box else 'off'}")
        g.set_mc_box(box)
        st = g.mc_run(prob.packets, seed=5, stream=1)
        res.append((st, g.download(lib.FIELD_TALLY)))
    (s0, t0), (s1, t1) = res
    assert s0.box_adds == 0 and s1.box_adds == 0
    for k in ("absorbed", "specular", "missed", "crossings", "interactions", "killed", "tally_adds"):
        assert getattr(s1, k) == getattr(s0, k), k
    np.testing.assert_allclose(t1, t0, rtol=1e-9, atol=1e-12)
    _print("case: 300 000 packets, 4 blocks per CU, box on")
    st = g.mc_run(300_000, seed=5, stream=1)
    assert st.box_adds > 0
    g.close()


def test_fixed_point_box_matches_fp64_box_at_c1_scale():
    """The 16-bit fixed-point LDS tally box against the FP64 one on the
    three-layer skin at 200^3 with 1e8 packets (the transport of one c1
    cycle), with the box placed from a previous run: the same packets and
    fates, and tallies that differ only by the rounding of each deposit to the
    box's unit (2^-13 of dx + dy + dz; at most half a unit per deposit,
    unbiased).  Every slot here wraps many times, so the carries into the
    tally are exercised.  Asserted: 1e-4 relative on every voxel above 1e-3 of
    the peak, 1e-6 on the total.  Two fixed-point runs give the same tally up
    to FP64 summation order: the integer sums do not depend on the order of
    the adds."""
    prob = configs.config1(n=(200, 200, 200), packets=100_000_000, steps=1, cycles=1)
    g = prob.grid()
    g.set_mc_launch(4, 32)
    res = []
    for fp64 in (True, False, False):
        _print(f"case: 200^3 skin, 1e8 packets, {'FP64' if fp64 else 'fixed-point'} box")
        g.set_mc_box(True, fp64=True)
        g.mc_run(prob.packets, seed=98, stream=0)          # the path each box is placed from
        g.set_mc_box(True, fp64=fp64)
        st = g.mc_run(prob.packets, seed=99, stream=1)
        res.append((st, g.download(lib.FIELD_TALLY), g.mc_box()))
    g.close()
    _print("boxes:", [r[2] for r in res])
    assert res[1][2] == res[2][2]                         # both fixed-point runs placed the same box
    res = [(st, t) for st, t, _ in res]
    (s64, t64), (sfx, tfx), (sfx2, tfx2) = res
    for k in ("absorbed", "specular", "missed", "crossings", "interactions", "killed", "tally_adds"):
        assert getattr(sfx, k) == getattr(s64, k), k
    _print("box adds per packet: FP64", s64.box_adds / prob.packets, "fixed point", sfx.box_adds / prob.packets)
    assert sfx.box_adds > 1.1 * s64.box_adds                # the 16-bit box holds four times the voxels
    hot = t64 > 1e-3 * t64.max()
    rel = np.abs(tfx[hot] - t64[hot]) / t64[hot]
    tot = abs(tfx.sum() / t64.sum() - 1)
    _print(f"voxels compared {int(hot.sum())}: max rel {rel.max():.3e}, median {np.median(rel):.3e}; "
           f"total rel {tot:.3e}")
    assert rel.max() < 1e-4
    assert tot < 1e-6
    np.testing.assert_allclose(tfx2, tfx, rtol=1e-12, atol=0)


def test_mc_box_follows_the_deposition_below_a_crater():
    """The LDS tally box takes the shape and z range that held the most of
    the previous run's path.  Below a deep crater that is the crater floor;
    box on and off still give the same fates and the same tally up to FP64
    summation order, and the oracle's fates.  The case is a 40^3 two-layer
    grid with a 24-plane crater: one run with the box off, then runs that
    place the box from the run before."""
    prob = _two_layer_problem(n=40, packets=200_000)
    n = 40
    mua = np.array([t[0] + t[1] for t in prob.table])
    rk = np.zeros((n + 2,) * 3)
    rk[1:-1, 1:-1, 1:-1] = mua[prob.mat]
    zz, yy, xx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    crater = ((xx - n / 2 + 0.5) ** 2 + (yy - n / 2 + 0.5) ** 2 <= 49) & (zz >= n - 24)
    rk[1:-1, 1:-1, 1:-1][crater] = 0.0
    prob.rhokap = rk
    g = prob.grid()
    g.set_mc_box(0)
    g.mc_run(prob.packets, seed=11, stream=0)
    jm = g.download(lib.FIELD_TALLY)
    assert jm[: n - 24].sum() > 0.3 * jm.sum()          # much of the path below the crater floor
    res = []
    for box in (1, 0, 1):
        _print(f"case: 40^3 crater 24 planes deep, FP64 box {box} placed from the previous run's path")
        g.set_mc_box(box, fp64=True)
        st = g.mc_run(prob.packets, seed=4321)
        res.append((st, g.download(lib.FIELD_TALLY)))
    g.close()
    st0, t0 = res[0]
    for st, tl in res[1:]:
        for k in ("absorbed", "specular", "missed", "crossings", "interactions", "killed", "tally_adds"):
            assert getattr(st, k) == getattr(st0, k), k
        np.testing.assert_allclose(tl, t0, rtol=1e-9, atol=1e-12)
    assert not t0[crater].any()
    t_or, ost = _oracle_for(prob, 4321).run(prob.packets)
    fg, fo = st0.fractions(), _fractions_oracle(ost, prob.packets)
    for k in fg:
        assert abs(fg[k] - fo[k]) <= 1e-3, k
    assert np.abs(t0 - t_or).sum() / t_or.sum() < 1e-2


def test_mc_box_placement_is_the_best_window():
    """The box's z range is the window of planes whose path (the tally the
    launch clears, over voxels with mu_t > 0, inside the box's footprint)
    is largest: recomputed in numpy from the downloaded tally and mu_t.
    Also: the footprint is centred on the beam, and a fresh grid (no path
    yet) gets the top planes."""
    prob = _two_layer_problem(n=40, packets=200_000)
    n = 40
    mua = np.array([t[0] + t[1] for t in prob.table])
    rk = np.zeros((n + 2,) * 3)
    rk[1:-1, 1:-1, 1:-1] = mua[prob.mat]
    zz, yy, xx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    crater = ((xx - n / 2 + 0.5) ** 2 + (yy - n / 2 + 0.5) ** 2 <= 30) & (zz >= n - 15)
    rk[1:-1, 1:-1, 1:-1][crater] = 0.0
    prob.rhokap = rk
    g = prob.grid()
    g.set_mc_launch(4, 32)
    g.set_mc_box(1, fp64=True)                                 # 27^3: the top-face window can move down
    g.mc_run(prob.packets, seed=5, stream=0)
    (x0, y0, z0), (sx, sy, sz) = g.mc_box()
    _print("fresh grid: box", (x0, y0, z0), (sx, sy, sz))
    assert sx == sy and z0 == n - sz                           # top planes without a path
    assert abs((x0 + sx / 2) - n / 2) <= 1 and abs((y0 + sy / 2) - n / 2) <= 1
    tally = g.download(lib.FIELD_TALLY)                        # the path the next launch places from
    mut = g.download(lib.FIELD_RHOKAP)[1:-1, 1:-1, 1:-1]
    g.mc_run(prob.packets, seed=6, stream=0)
    (x1, y1, z1), n1 = g.mc_box()
    g.close()
    assert (x1, y1) == (x0, y0) and n1 == (sx, sy, sz)
    foot = np.where(mut > 0, tally, 0.0)[:, y0:y0 + sy, x0:x0 + sx].sum(axis=(1, 2))
    win = np.array([foot[o:o + sz].sum() for o in range(n - sz + 1)])
    best = win.max()
    _print("placed at z", z1, "window sum", win[z1], "best", best, "at", int(np.flatnonzero(win == best)[-1]))
    assert z1 < n - sz                                         # moved below the crater floor
    assert win[z1] >= best * (1 - 1e-12)                       # the best window (to FP64 summation order)


def _batches(g, nb, per, seed):
    out = []
    for b in range(nb):
        g.mc_run(per, seed=seed, stream=b)
        out.append(g.download(lib.FIELD_TALLY))
    return np.array(out)


def test_mc_independent_streams_within_3_sigma():
    prob = _two_layer_problem(n=24)
    nb, per = 10, 40_000
    _print("case: two-layer 24^3, 10 x 4e4 packets per side, independent streams")
    g = prob.grid()
    gb = _batches(g, nb, per, seed=99)
    g.close()
    ob = []
    for b in range(nb):
        t, _ = _oracle_for(prob, 5, stream=b).run(per)
        ob.append(t)
    ob = np.array(ob)
    mg, mo = gb.mean(0), ob.mean(0)
    se = np.sqrt(gb.var(0, ddof=1) / nb + ob.var(0, ddof=1) / nb)
    ok = (mg > 0) & (mo > 0) & (se > 0)
    z = np.abs(mg - mo)[ok] / se[ok]
    frac = (z <= 3).mean()
    _print(f"voxels tested {ok.sum()}, within 3 sigma {frac:.4f}")
    assert frac >= 0.99


def test_mc_reference_physics_beer_lambert():
    """Reference default physics (pure absorber, mu_a = 680/cm, 125 um top-hat):
    per-layer tally = exp(-mu z1) - exp(-mu z2) per packet (closed form)."""
    prob = configs.reference_default()
    g = prob.grid()
    npk = 2_000_000
    st = g.mc_run(npk, seed=3)
    t = g.download(lib.FIELD_TALLY)
    g.close()
    p = prob.params
    dz = 2 * p.zmax / p.nz
    mu = 680.0
    layer = (t * mu).sum(axis=(1, 2))[::-1] / npk        # from the top down
    z = np.arange(p.nz + 1) * dz
    expect = np.exp(-mu * z[:-1]) - np.exp(-mu * z[1:])
    # variance of a layer sum is below the Bernoulli bound expect*(1-expect)/npk
    sig = np.sqrt(expect * (1 - expect) / npk) + 1e-12
    sel = expect > 1e-4
    zsc = np.abs(layer - expect)[sel] / sig[sel]
    _print("layers", sel.sum(), "max z", zsc.max())
    assert zsc.max() < 5
    assert st.absorbed == npk


def test_mc_reference_physics_vs_ran2_walk():
    """GPU (Philox, FP32 DDA walk) vs the reference's own walk (ran2, FP64,
    bisection + nudge) restated in the oracle: per-layer sums within 4 sigma."""
    prob = configs.reference_default()
    g = prob.grid()
    npk = 400_000
    g.mc_run(npk, seed=11)
    t = g.download(lib.FIELD_TALLY) * 680.0
    g.close()
    jo, esc = O.ref_mc_oracle(80, 0.03, 0.03, 0.06, npk, -95648324)
    lg, lo = t.sum(axis=(1, 2)), jo.sum(axis=(1, 2))
    sig = np.sqrt(2 * np.maximum(lo, 1) / npk) * np.sqrt(npk)   # Poisson-like bound per layer
    assert np.all(np.abs(lg - lo) <= 4 * sig + 1e-9)


# ---------------------------------------------------------------- heat ----

def _oracle_heat(prob, T, kappa, density, heatcap, jmean, steps, Q=None, laser=None):
    p = prob.params
    nx, ny, nz = p.nx, p.ny, p.nz
    coeff = np.zeros_like(T)
    coeff[1:-1, 1:-1, 1:-1] = p.delt / (density * heatcap)[1:-1, 1:-1, 1:-1]
    Tt = np.ascontiguousarray(T.copy())
    Qa = np.zeros((nz, ny, nx)) if Q is None else np.ascontiguousarray(Q.copy())
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (kappa, density, heatcap, coeff, jmean)]
    h = O.Heat()
    h.nx, h.ny, h.nz = nx, ny, nz
    h.dx, h.dy, h.dz = p.dx_m, p.dy_m, p.dz_m
    h.delt, h.dscale = p.delt, p.dscale
    h.temp = Tt.ctypes.data
    h.kappa, h.density, h.heatcap, h.coeff, h.jmean = [a.ctypes.data for a in arrs]
    h.Q = Qa.ctypes.data
    h.volumeVoxel, h.massVoxel, h.QVapor = p.volume_voxel, p.mass_voxel, p.qvapor
    h.boil, h.robin_top, h.h_top, h.T_amb = p.boil, p.robin_top, p.h_top, p.t_amb
    lon = None if laser is None else np.ascontiguousarray(laser, dtype=np.float64)
    rc = O.lib().or_heat_steps(C.byref(h), steps, None if lon is None else lon.ctypes.data)
    assert rc == 0
    return Tt, Qa, h.flux_out


def _random_thermal(prob, rng, spread=0.2):
    p = prob.params
    hs = (p.nz + 2, p.ny + 2, p.nx + 2)
    f = lambda v: v * (1 + spread * (rng.random(hs) - 0.5))
    return f(prob.kappa), f(prob.density), f(prob.heatcap)


def test_heat_config_mode_robin_within_1e5():
    prob = configs.config0(n=30, packets=1, steps=600)
    rng = np.random.default_rng(1)
    kappa, density, heatcap = _random_thermal(prob, rng)
    p = prob.params
    nz, ny, nx = 30, 30, 30
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    jmean = 4e6 * np.exp(-((xx - 15) ** 2 + (yy - 15) ** 2) / 20.0 - (nz - 1 - zz) / 6.0)
    T = prob.temp.copy()
    _print("case: config-mode heat 30^3, 600 steps, random properties, Robin top")
    g = lib.TissueGrid(p)
    g.set_optics(prob.mat, prob.table)
    g.set_thermal(T, kappa, density, heatcap)
    g.upload(lib.FIELD_JMEAN, jmean)
    g.heat_run(prob.steps)
    Tg = g.download(lib.FIELD_TEMP)
    hc_g, out_g = g.energy()
    g.close()
    To, _, flux = _oracle_heat(prob, T, kappa, density, heatcap, jmean, prob.steps)
    inner = (slice(1, -1),) * 3
    rel = np.abs(Tg[inner] - To[inner]) / To[inner]
    _print("max rel", rel.max(), "max rise", (To[inner] - p.t_ref).max())
    assert rel.max() <= 1e-5
    assert abs(out_g - flux) <= 1e-4 * abs(flux) + 1e-9


def test_heat_reference_mode_boiling_within_1e5():
    prob = configs.reference_default(n=24)
    p = prob.params
    rng = np.random.default_rng(2)
    T = prob.temp.copy()
    T[1:-1, 1:-1, 1:-1] = 360 + 20 * rng.random((24, 24, 24))        # some voxels boil
    jmean = 5e9 * rng.random((24, 24, 24))
    _print("case: reference-mode heat 24^3, 200 steps, boiling on")
    g = prob.grid()
    g.upload(lib.FIELD_TEMP, T)
    g.upload(lib.FIELD_JMEAN, jmean)
    g.heat_run(200)
    Tg = g.download(lib.FIELD_TEMP)
    Qg = g.download(lib.FIELD_Q)
    g.close()
    To, Qo, _ = _oracle_heat(prob, T, prob.kappa, prob.density, prob.heatcap, jmean, 200)
    inner = (slice(1, -1),) * 3
    rel = np.abs(Tg[inner] - To[inner]) / To[inner]
    _print("max rel T", rel.max(), "boiling voxels", (Qo > 0).sum(),
           "bit-identical voxels", (Tg == To).mean())
    assert (Qo > 0).sum() > 0
    # FP64 path, reference operation order, no FMA: identical to the oracle
    assert np.array_equal(Tg, To)
    assert np.array_equal(Qg, Qo)


def test_energy_balance_fp64():
    """Heat content change = deposited source energy + boundary inflow (uniform rho c)."""
    prob = configs.config0(n=24, packets=1, steps=300)
    p = prob.params
    jmean = np.zeros((24, 24, 24))
    jmean[8:16, 8:16, 8:16] = 1e7
    g = prob.grid()
    g.upload(lib.FIELD_JMEAN, jmean)
    hc0, _ = g.energy()
    g.heat_run(prob.steps)
    hc1, out = g.energy()
    g.close()
    src = jmean.sum() * p.volume_voxel * p.delt * prob.steps         # J
    err = (hc1 - hc0) - (src - out)
    _print("heat content change", hc1 - hc0, "source", src, "boundary out", out, "err", err)
    assert abs(err) <= 2e-5 * src


# ------------------------------------------------- Arrhenius and rebuild --

def _arrh(p):
    a = O.Arrh()
    a.A, a.dE, a.R, a.Tlo, a.Thi = p.arr_A, p.arr_dE, p.arr_R, p.arr_Tlo, p.arr_Thi
    a.thr = (C.c_double * 3)(*p.arr_thr)
    return a


def test_arrhenius_reference_cadence():
    prob = configs.reference_default(n=20)
    p = prob.params
    rng = np.random.default_rng(3)
    T = prob.temp.copy()
    T[1:-1, 1:-1, 1:-1] = 330 + 40 * rng.random((20, 20, 20))
    g = prob.grid()
    g.upload(lib.FIELD_TEMP, T)
    Tdev = g.download(lib.FIELD_TEMP)
    for it in range(3):
        g.arrhenius(0.3, 0.5 * (it + 1))
    om_g = g.download(lib.FIELD_TISSUE)
    th_g = g.download(lib.FIELD_THRESTIME)
    g.close()
    om = np.zeros((20, 20, 20))
    th = np.zeros((3, 20, 20, 20))
    for it in range(3):
        O.lib().or_arrhenius(C.byref(_arrh(p)), 20, 20, 20, np.ascontiguousarray(Tdev),
                             np.ascontiguousarray(prob.rhokap), 0.3, 0.5 * (it + 1), om, th)
    assert np.allclose(om_g, om, rtol=1e-12, atol=0)
    assert np.array_equal(th_g, th)
    assert (th[0] > 0).any() and (th[1] > 0).any()


@pytest.mark.parametrize("shape", [None, 8])         # 8: two steps per launch
def test_fused_damage_constant_temperature_closed_form(shape):
    """FP32 path with Omega integrated inside every step: a uniform
    temperature in the window (Dirichlet halo at the same value, no source)
    stays put, so Omega grows by d = delt A exp(-Ea/RT) per step and the
    threshold times follow in closed form, including the chain's one time per
    step (0.53 and 1 are both passed in step 1; the second time is written in
    step 2).  d is chosen so that 1e4 / d is half-way between two steps."""
    import math
    p = configs._config_params((12, 10, 9), (0.12, 0.10, 0.09))
    p.robin_top = 0
    p.boil = 0
    p.damage_every_step = 1
    d = p.arr_thr[2] / 200.5
    T_star = p.arr_dE / (p.arr_R * (math.log(p.arr_A) + math.log(p.delt) - math.log(d)))
    assert p.arr_Tlo < T_star < p.arr_Thi
    mat = np.zeros((9, 10, 12), np.uint8)
    temp, kappa, density, heatcap = configs._thermal_cube(p, configs.C0_THERMAL, t0=T_star)
    prob = configs.Problem("uniform-window", p, mat, [(0.5, 100.0, 0.9, 1.38)], temp, kappa, density, heatcap)
    g = prob.grid()
    if shape is not None:
        g.set_heat_launch(shape, 4)
    steps = 205
    g.heat_run(steps)
    T = g.download(lib.FIELD_TEMP)
    om = g.download(lib.FIELD_TISSUE)
    th = g.download(lib.FIELD_THRESTIME)
    g.close()
    T_dev = T[1:-1, 1:-1, 1:-1]
    assert np.ptp(T_dev) == 0.0                           # uniform and unchanged
    d_dev = p.delt * p.arr_A * np.exp(-p.arr_dE / (p.arr_R * T_dev[0, 0, 0]))
    np.testing.assert_allclose(om, steps * d_dev, rtol=2e-5)
    t = np.cumsum(np.full(steps, p.delt))                 # time after each step
    np.testing.assert_allclose(th[0], t[0], rtol=1e-12)
    np.testing.assert_allclose(th[1], t[1], rtol=1e-12)
    np.testing.assert_allclose(th[2], t[200], rtol=1e-12)


def test_arrhenius_chain_with_uploaded_threshold_times():
    """The threshold chain's "ThresTime(k) == 0" tests come from a per-voxel
    mask rebuilt on upload: uploaded times with gaps (k = 1 set, 2 clear, 3
    set, and every other pattern) must continue exactly as the oracle's chain
    on the FP64 times."""
    prob = configs.reference_default(n=16)
    p = prob.params
    rng = np.random.default_rng(11)
    T = prob.temp.copy()
    T[1:-1, 1:-1, 1:-1] = 340 + 30 * rng.random((16, 16, 16))
    th0 = np.zeros((3, 16, 16, 16))
    pattern = rng.integers(0, 8, size=(16, 16, 16))
    for k in range(3):
        th0[k][(pattern >> k) & 1 == 1] = 0.25 * (k + 1)
    om0 = 2e4 * rng.random((16, 16, 16)) ** 4            # every threshold reachable
    g = prob.grid()
    g.upload(lib.FIELD_TEMP, T)
    g.upload(lib.FIELD_TISSUE, om0)
    g.upload(lib.FIELD_THRESTIME, th0)
    Tdev = g.download(lib.FIELD_TEMP)
    for it in range(4):
        g.arrhenius(0.2, 1.0 + it)
    om_g = g.download(lib.FIELD_TISSUE)
    th_g = g.download(lib.FIELD_THRESTIME)
    g.close()
    om, th = om0.copy(), th0.copy()
    for it in range(4):
        O.lib().or_arrhenius(C.byref(_arrh(p)), 16, 16, 16, np.ascontiguousarray(Tdev),
                             np.ascontiguousarray(prob.rhokap), 0.2, 1.0 + it, om, th)
    assert np.allclose(om_g, om, rtol=1e-12, atol=0)
    assert np.array_equal(th_g, th)
    assert (th != th0).sum() > 100                        # the chain fired, including past gaps


def _rebuild_oracle(prob, Tdev, rhokap, mode, Q=None, W=None):
    p = prob.params
    nz, ny, nx = p.nz, p.ny, p.nx
    arrs = dict(rhokap=rhokap.copy(), density=prob.density.copy(), heatcap=prob.heatcap.copy(),
                kappa=prob.kappa.copy(), coeff=np.zeros_like(Tdev), alpha=np.zeros_like(Tdev))
    arrs["coeff"][1:-1, 1:-1, 1:-1] = p.delt / (prob.density * prob.heatcap)[1:-1, 1:-1, 1:-1]
    arrs = {k: np.ascontiguousarray(v) for k, v in arrs.items()}
    Wa = np.ascontiguousarray(np.full((nz, ny, nx), 0.75) if W is None else W.copy())
    Qa = np.ascontiguousarray(np.zeros((nz, ny, nx)) if Q is None else Q)
    Ta = np.ascontiguousarray(Tdev)
    r = O.Rebuild()
    r.nx, r.ny, r.nz, r.mode = nx, ny, nz, mode
    r.ablate_K, r.delt = p.ablate_K, p.delt
    r.mu_water, r.mu_protein, r.QVapor = p.mu_water, p.mu_protein, p.qvapor
    for k in ("rhokap", "density", "heatcap", "kappa", "coeff", "alpha"):
        setattr(r, k, arrs[k].ctypes.data)
    r.W, r.Q, r.temp = Wa.ctypes.data, Qa.ctypes.data, Ta.ctypes.data
    O.lib().or_rebuild(C.byref(r))
    arrs["W"] = Wa
    return arrs


@pytest.mark.parametrize("mode", [lib.THERMAL_WATER, lib.THERMAL_TABLE])
def test_rebuild_mask_bitexact(mode):
    n = 20
    prob = configs.reference_default(n=n)
    p = prob.params
    p.thermal_mode = mode
    p.ablate_K = 400.0
    rng = np.random.default_rng(10 + mode)
    rk = prob.rhokap.copy()
    holes = rng.random((n, n, n)) < 0.45                      # many empty voxels
    rk[1:-1, 1:-1, 1:-1][holes] = 0.0
    prob.rhokap = rk
    T = prob.temp.copy()
    T[1:-1, 1:-1, 1:-1] = 380 + 30 * rng.random((n, n, n))   # ~2/3 below the ablation point
    Q = rng.random((n, n, n)) * p.qvapor * 0.5
    g = prob.grid()
    g.upload(lib.FIELD_TEMP, T)
    g.upload(lib.FIELD_Q, Q)
    Tdev = g.download(lib.FIELD_TEMP)
    g.rebuild()
    out = {f: g.download(fid) for f, fid in (("rhokap", lib.FIELD_RHOKAP), ("density", lib.FIELD_DENSITY),
                                           ("heatcap", lib.FIELD_HEATCAP), ("kappa", lib.FIELD_KAPPA),
                                           ("coeff", lib.FIELD_COEFF), ("W", lib.FIELD_WATER))}
    g.close()
    ref = _rebuild_oracle(prob, Tdev, rk, mode, Q=Q)
    zero_g = out["rhokap"][1:-1, 1:-1, 1:-1] == 0
    zero_o = ref["rhokap"][1:-1, 1:-1, 1:-1] == 0
    _print("removed", zero_o.sum(), "of", n ** 3)
    assert np.array_equal(zero_g, zero_o)                      # the mask: bit-exact
    assert np.array_equal(out["rhokap"], ref["rhokap"])
    for k in ("density", "heatcap", "coeff", "W"):
        assert np.array_equal(out[k], ref[k]), k
    assert np.allclose(out["kappa"], ref["kappa"], rtol=1e-15, atol=0)   # exp() may differ by 1 ulp

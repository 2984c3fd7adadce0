"""Whole coupling cycles of the HIP path against the oracle, cycle by cycle
(mcpolar.f90:148-186): transport -> normalisation -> heat steps (with the
damage integral) -> rebuild, the oracle fed the same inputs as the GPU at the
start of each stage (so differences do not compound):

  * the three-layer-skin workload of the 200^3 benchmark at 40^3 (Gaussian
    beam, FP32 stencil with FP64 Arrhenius inside every step, ablation and
    rebuild), over cycles that ablate;
  * the reference's own physics (FP64 stencil with boiling, water model,
    Arrhenius at the reference cadence, the driver's pulse and scale), whose
    final fields are compared through the writer's eight arrays
    (writer.f90:21-79, mcpolar.f90:190-201).

Tolerances: transport on the same Philox streams, fate fractions 1e-3
absolute and tally L1 1e-2 (FP32 walk vs FP64); jmean bit-exact against
numpy on the GPU tally; FP32 temperatures 1e-5 relative (FP64 path: equal);
the damage integral 1e-4 relative (expf of the FP64 exponent), threshold
times equal except where the FP32 and FP64 fields cross a threshold in
adjacent steps (at most 0.1% of the voxels, one step apart); the ablation
mask and rhokap bit-exact given the GPU's temperature.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_ffi as O
from tam_amd import configs, driver, lib

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _table_with_air(prob):
    return list(prob.table) + [(0.0, 0.0, 0.0, 1.0)]


def _oracle_mc(prob, g, seed, stream):
    p = prob.params
    mat = g.download(lib.FIELD_MAT).astype(np.uint8)
    rk = g.download(lib.FIELD_RHOKAP)[1:-1, 1:-1, 1:-1]
    m = O.MCProblem(p.nx, p.ny, p.nz, p.xmax, p.ymax, p.zmax, mat, _table_with_air(prob), mut=rk,
                    n_ext=(0, 0, 0, 0, p.n_ext_zlo, p.n_ext_zhi), beam=p.beam, radius=p.beam_radius,
                    rng=0, seed=seed, stream=stream)
    return m.run(prob.packets)


def _check_transport(prob, st, t_gpu, ost, t_or):
    n = prob.packets
    fg = st.fractions()
    fo = {"absorbed": ost.absorbed / n, "specular": ost.specular / n, "diffuse_reflected": ost.escaped[5] / n,
          "transmitted": ost.escaped[4] / n, "lateral": sum(ost.escaped[:4]) / n, "missed": ost.missed / n}
    for k in fg:
        assert abs(fg[k] - fo[k]) <= 1e-3, (k, fg[k], fo[k])
    rel = np.abs(t_gpu - t_or).sum() / t_or.sum()
    assert rel < 1e-2, rel
    return rel


def _check_normalise(prob, g, tally, scale):
    mat = g.download(lib.FIELD_MAT).astype(np.int64)
    rk = g.download(lib.FIELD_RHOKAP)[1:-1, 1:-1, 1:-1]
    frac = np.array([t[0] / (t[0] + t[1]) if t[0] + t[1] > 0 else 0.0 for t in _table_with_air(prob)])
    jm = g.download(lib.FIELD_JMEAN)
    assert np.array_equal(jm, tally * (rk * frac[mat]) * scale)
    return jm


class _OracleHeat:
    """or_heat_steps on the GPU's downloaded state, one step per call."""

    def __init__(self, prob, g, jm):
        p = prob.params
        self.p = p
        self.T = np.ascontiguousarray(g.download(lib.FIELD_TEMP))
        self.arrs = [np.ascontiguousarray(g.download(f)) for f in
                     (lib.FIELD_KAPPA, lib.FIELD_DENSITY, lib.FIELD_HEATCAP, lib.FIELD_COEFF, lib.FIELD_RHOKAP)]
        self.jm = np.ascontiguousarray(jm)
        self.Q = np.ascontiguousarray(g.download(lib.FIELD_Q))
        h = O.Heat()
        h.nx, h.ny, h.nz = p.nx, p.ny, p.nz
        h.dx, h.dy, h.dz = p.dx_m, p.dy_m, p.dz_m
        h.delt, h.dscale = p.delt, p.dscale
        h.temp = self.T.ctypes.data
        h.kappa, h.density, h.heatcap, h.coeff, h.rhokap = [a.ctypes.data for a in self.arrs]
        h.jmean, h.Q = self.jm.ctypes.data, self.Q.ctypes.data
        h.volumeVoxel, h.massVoxel, h.QVapor = p.volume_voxel, p.mass_voxel, p.qvapor
        h.boil, h.robin_top, h.h_top, h.T_amb = p.boil, p.robin_top, p.h_top, p.t_amb
        h.removed_inert = p.removed_inert
        self.h = h

    def step(self, laser_on=1.0):
        lon = np.array([laser_on])
        assert O.lib().or_heat_steps(C.byref(self.h), 1, lon.ctypes.data) == 0


def _arrh(p):
    a = O.Arrh()
    a.A, a.dE, a.R, a.Tlo, a.Thi = p.arr_A, p.arr_dE, p.arr_R, p.arr_Tlo, p.arr_Thi
    a.thr = (C.c_double * 3)(*p.arr_thr)
    return a


def _rebuild_oracle(prob, g, T):
    """or_rebuild on the GPU's state before its rebuild, given the GPU's T."""
    p = prob.params
    arrs = {k: np.ascontiguousarray(g.download(f)) for k, f in
            (("rhokap", lib.FIELD_RHOKAP), ("density", lib.FIELD_DENSITY), ("heatcap", lib.FIELD_HEATCAP),
             ("kappa", lib.FIELD_KAPPA), ("coeff", lib.FIELD_COEFF))}
    arrs["alpha"] = np.zeros_like(arrs["rhokap"])
    W = np.ascontiguousarray(g.download(lib.FIELD_WATER))
    Q = np.ascontiguousarray(g.download(lib.FIELD_Q))
    Ta = np.ascontiguousarray(T)
    r = O.Rebuild()
    r.nx, r.ny, r.nz, r.mode = p.nx, p.ny, p.nz, p.thermal_mode
    r.ablate_K, r.delt = p.ablate_K, p.delt
    r.mu_water, r.mu_protein, r.QVapor = p.mu_water, p.mu_protein, p.qvapor
    for k in ("rhokap", "density", "heatcap", "kappa", "coeff", "alpha"):
        setattr(r, k, arrs[k].ctypes.data)
    r.W, r.Q, r.temp = W.ctypes.data, Q.ctypes.data, Ta.ctypes.data
    O.lib().or_rebuild(C.byref(r))
    arrs["W"] = W
    return arrs


def _check_damage(p, om_g, om_o, th_g, th_o, Tg):
    """Omega to 1e-4 relative; where the FP32 and FP64 fields put a voxel on
    different sides of the window edge or a threshold in one step, at most
    0.1% of voxels differ, by at most two steps' damage / one step's time."""
    big = np.maximum(np.abs(om_o), 1e-300)
    bad = np.abs(om_g - om_o) > 1e-4 * big
    step_max = p.delt * p.arr_A * math.exp(-p.arr_dE / (p.arr_R * p.arr_Thi))
    assert bad.mean() <= 1e-3, bad.mean()
    assert np.all(np.abs(om_g - om_o)[bad] <= 2 * step_max)
    dth = th_g != th_o
    assert dth.mean() <= 1e-3, dth.mean()
    assert np.all(np.abs(th_g - th_o)[dth] <= 1.0001 * p.delt)


def test_skin_cycles_follow_the_oracle_chain():
    """40^3 three-layer skin (Gaussian beam, 40 W so that it ablates), 4
    cycles of 1.5e5 packets and 60 steps, stepwise through the C-ABI."""
    prob = configs.config1(n=(40, 40, 40), packets=150_000, steps=60, cycles=4)
    prob.power = 40.0
    p = prob.params
    g = prob.grid()
    t = 0.0                                   # the library's own clock: time += delt per step
    arr = _arrh(p)
    for cyc in range(4):
        _print(f"case: 40^3 skin chain, cycle {cyc}")
        st = g.mc_run(prob.packets, seed=7, stream=cyc)
        tally = g.download(lib.FIELD_TALLY)
        t_or, ost = _oracle_mc(prob, g, 7, cyc)
        rel = _check_transport(prob, st, tally, ost, t_or)
        g.normalise(prob.scale)
        jm = _check_normalise(prob, g, tally, prob.scale)
        oh = _OracleHeat(prob, g, jm)
        om = np.ascontiguousarray(g.download(lib.FIELD_TISSUE))
        th = np.ascontiguousarray(g.download(lib.FIELD_THRESTIME))
        rk = oh.arrs[4]
        g.heat_run(prob.steps)
        for s in range(prob.steps):
            oh.step()
            ta = t + p.delt
            O.lib().or_arrhenius(C.byref(arr), p.nx, p.ny, p.nz, oh.T, rk, p.delt, ta, om, th)
            t += p.delt
        Tg = g.download(lib.FIELD_TEMP)
        inner = (slice(1, -1),) * 3
        relT = (np.abs(Tg - oh.T)[inner] / oh.T[inner]).max()
        _check_damage(p, g.download(lib.FIELD_TISSUE), om, g.download(lib.FIELD_THRESTIME), th, Tg)
        ref = _rebuild_oracle(prob, g, Tg)
        g.rebuild()
        rk_g = g.download(lib.FIELD_RHOKAP)
        removed = int((rk_g[inner] == 0).sum())
        _print(f"cycle {cyc}: tally L1 {rel:.2e}, max rel T {relT:.2e}, max T {Tg.max():.1f} K, "
               f"omega>=1 {(om >= 1).sum()}, removed {removed}")
        assert relT <= 1e-5
        assert np.array_equal(rk_g == 0, ref["rhokap"] == 0)
        assert np.array_equal(rk_g, ref["rhokap"])
    assert removed > 0, "the case should ablate"
    assert (om >= 1.0).any()
    g.close()


def test_reference_mode_chain_and_writer_fields():
    """The reference default (input.params, 24^3, tophat pulse of 30 steps,
    loops = 1, FP64 stencil with boiling, Arrhenius at the reference cadence,
    water-model rebuild) for 40 cycles, stepwise: each stage against the
    oracle on the GPU's state (FP64 heat bit-identical), then the writer's
    eight fields against the oracle chain's."""
    inp = dict(configs.REFERENCE_INPUT, pulsetype="tophat", nphotons=20_000)
    n = 24
    prob = configs.reference_default(inp, n)
    p = prob.params
    prob.packets = inp["nphotons"]
    g = prob.grid()
    pulse = lib.Pulse()
    pulse.realPulseLength = 30 * p.delt
    pulse.repetitionRate_1 = 1e30
    pulse.laserOn, pulse.laser_flag, pulse.pulsesToDo, pulse.use_pulse = 1.0, 1, 1, 1
    power = driver.PowerFunction(inp, prob.pulse_length)
    arr = _arrh(p)
    for cyc in range(40):
        on = pulse.laser_flag != 0
        if on:
            st = g.mc_run(prob.packets, seed=driver.SEED, stream=cyc)
            tally = g.download(lib.FIELD_TALLY)
            t_or, ost = _oracle_mc(prob, g, driver.SEED, cyc)
            _check_transport(prob, st, tally, ost, t_or)
            # 3e-3 of the input's 70 W on the whole grid: the top planes
            # reach the boiling plateau and the damage thresholds within the
            # pulse without being ablated (the full power vaporises them in a
            # few steps)
            scale = 3e-3 * power(pulse) / (prob.packets * p.volume_voxel)
            g.normalise(scale)
            jm = _check_normalise(prob, g, tally, scale)
        else:
            jm = g.download(lib.FIELD_JMEAN)          # laser off: jmeanGLOBAL is kept
        oh = _OracleHeat(prob, g, jm)
        lon = pulse.laserOn
        g.heat_run(1, pulse)
        oh.step(lon)
        Tg = g.download(lib.FIELD_TEMP)
        assert np.array_equal(Tg, oh.T), cyc
        assert np.array_equal(g.download(lib.FIELD_Q), oh.Q), cyc
        om = np.ascontiguousarray(g.download(lib.FIELD_TISSUE))
        th = np.ascontiguousarray(g.download(lib.FIELD_THRESTIME))
        rk = np.ascontiguousarray(g.download(lib.FIELD_RHOKAP))
        g.arrhenius(p.delt, pulse.time)
        O.lib().or_arrhenius(C.byref(arr), n, n, n, np.ascontiguousarray(Tg), rk, p.delt, pulse.time, om, th)
        om_g, th_g = g.download(lib.FIELD_TISSUE), g.download(lib.FIELD_THRESTIME)
        np.testing.assert_allclose(om_g, om, rtol=1e-12, atol=0)
        assert np.array_equal(th_g, th), cyc
        ref = _rebuild_oracle(prob, g, Tg)
        g.rebuild()
        assert np.array_equal(g.download(lib.FIELD_RHOKAP), ref["rhokap"]), cyc
        for k, f in (("density", lib.FIELD_DENSITY), ("heatcap", lib.FIELD_HEATCAP), ("coeff", lib.FIELD_COEFF),
                     ("W", lib.FIELD_WATER)):
            assert np.array_equal(g.download(f), ref[k]), (cyc, k)
        np.testing.assert_allclose(g.download(lib.FIELD_KAPPA), ref["kappa"], rtol=1e-15, atol=0)
    _print("max T", Tg.max(), "laser", pulse.laser_flag, "pulses done", pulse.pulsesDone)
    # the writer's fields (mcpolar.f90:190-201, writer.f90:21-79) from the
    # last oracle stage: temperature in C with halo, tissue = -1 in the
    # crater, the three threshold-time planes
    fields = driver.writer_fields(g)
    g.close()
    rk_i = ref["rhokap"][1:-1, 1:-1, 1:-1]
    tissue = om.copy()
    tissue[rk_i <= 0.1] = -1.0
    expect = {"jmean-t": jm, "rhokap-t": rk_i, "temp-t": Tg - 273.0, "water-t": ref["W"], "tissue-t": tissue,
              "/time-t-1-": th[0], "/time-t-2-": th[1], "/time-t-3-": th[2]}
    assert sorted(fields) == sorted(expect)
    for k, a in expect.items():
        if k in ("tissue-t",):
            np.testing.assert_allclose(fields[k], a, rtol=1e-12, atol=0, err_msg=k)
        else:
            assert np.array_equal(fields[k], a), k
    assert Tg.max() > 300.0 and (th[0] > 0).any()


@pytest.mark.parametrize("shape,first", [(None, 1), (8, 1), (8, 2), (4, 2)])
def test_threshold_chain_continues_outside_the_window(shape, first):
    """The reference's chain runs for every voxel on every call and writes at
    most one threshold time per call (3dFD.f90:455-461).  A uniform
    temperature in the window whose one-step damage passes all three
    thresholds writes the first time in that step; the field is then moved
    above the window (380 K, no more damage), and the later times are still
    written one step apart, as per-step calls of the reference's Arrhenius
    would write them.  `first` steps run before the move (so the due chain
    crosses a one-step or a two-step launch boundary)."""
    p = configs._config_params((12, 10, 9), (0.12, 0.10, 0.09))
    p.robin_top = 0
    p.boil = 0
    p.damage_every_step = 1
    d = p.arr_thr[2] / 0.6                              # one step passes 0.53, 1 and 1e4
    T_star = p.arr_dE / (p.arr_R * (math.log(p.arr_A) + math.log(p.delt) - math.log(d)))
    assert p.arr_Tlo < T_star < p.arr_Thi
    mat = np.zeros((9, 10, 12), np.uint8)
    temp, kappa, density, heatcap = configs._thermal_cube(p, configs.C0_THERMAL, t0=T_star)
    prob = configs.Problem("uniform-window", p, mat, [(0.5, 100.0, 0.9, 1.38)], temp, kappa, density, heatcap)
    g = prob.grid()
    if shape is not None:
        g.set_heat_launch(shape, 4)
    g.heat_run(first)
    g.upload(lib.FIELD_TEMP, np.full_like(temp, 380.0))
    g.heat_run(4)
    th = g.download(lib.FIELD_THRESTIME)
    om = g.download(lib.FIELD_TISSUE)
    g.close()
    t = np.cumsum(np.full(first + 4, p.delt))           # time after each step, as the library adds it
    np.testing.assert_allclose(om, first * p.delt * p.arr_A * np.exp(-p.arr_dE / (p.arr_R * T_star)), rtol=2e-5)
    assert np.all(th[0] == t[0]) and np.all(th[1] == t[1]) and np.all(th[2] == t[2])

"""Host-side logic (CPU only): input.params parsing, the reference-default
set-up (initThermalCoeff) and the benchmark configurations."""
import ctypes as C
import math
import os

import numpy as np

import oracle_ffi as O
from tam_amd import configs

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_input_params_parsing():
    """The reference's 14 positional lines (mcpolar.f90:79-94)."""
    inp = configs.read_input_params(os.path.join(GOLD, "input.params"))
    assert inp == configs.REFERENCE_INPUT


def test_reference_default_matches_init_thermal():
    prob = configs.reference_default()
    p = prob.params
    t = O.ThermalInit()
    O.lib().or_init_thermal(80, 0.03, 0.03, 0.06, 80, 80, 80, 400.0, 70.0, C.byref(t))
    for a, b in ((p.dx_m, t.dx), (p.dy_m, t.dy), (p.dz_m, t.dz), (p.delt, t.delt),
                 (p.volume_voxel, t.volumeVoxel), (p.mass_voxel, t.massVoxel), (p.qvapor, t.QVapor)):
        assert math.isclose(a, b, rel_tol=1e-15)
    assert math.isclose(prob.pulse_length, t.pulseLength, rel_tol=1e-15)
    assert p.dscale == 0.5 and p.boil == 1 and p.heat_fp64 == 1
    # halo conductivity is air at 298 K, density and heat capacity are skin (3dFD.f90:272-278)
    assert math.isclose(prob.kappa[0, 0, 0], t.kappa_air, rel_tol=1e-15)
    assert prob.density[0, 0, 0] == t.density and prob.heatcap[0, 0, 0] == t.heatcap
    # temperature initialisation (mcpolar.f90:123-129)
    assert prob.temp[0, 5, 5] == 298.0 and prob.temp[-1, 5, 5] == 298.0 and prob.temp[5, 0, 5] == 278.0


def test_explicit_limit_is_the_fourier_stability_bound():
    """At dt = limit the highest mode's amplification is exactly -1."""
    alpha, dx, dy, dz = 1.5e-7, 1e-4, 1e-4, 5e-5
    for s in (0.5, 1.0):
        dt = configs.explicit_limit(alpha, dx, dy, dz, s)
        lam = 1 - 4 * s * alpha * dt * (1 / dx ** 2 + 1 / dy ** 2 + 1 / dz ** 2)
        assert math.isclose(lam, -1.0, rel_tol=1e-12)


def test_three_layer_skin_layers():
    prob = configs.three_layer_skin(n=(20, 20, 200), packets=10, steps=1, cycles=1)
    mz = prob.mat[:, 0, 0][::-1]                 # from the top down
    # epidermis 0.01 cm = 2 planes of 0.005 cm, dermis 0.2 cm = 40 planes
    assert (mz[:2] == 0).all() and (mz[2:42] == 1).all() and (mz[42:] == 2).all()
    assert prob.params.beam == 1 and prob.params.beam_radius == 0.15
    assert prob.params.ablate_K == 373.0
    assert math.isclose(prob.scale * prob.packets * prob.params.volume_voxel, prob.power)


def test_config0_thermal_and_time_step():
    prob = configs.config0(n=100)
    p = prob.params
    alpha = 0.56 / (1070 * 3400)
    assert math.isclose(p.delt, 0.9 * (1e-4) ** 2 / (6 * alpha), rel_tol=1e-12)   # 9.74e-3 s
    assert math.isclose(p.delt, 9.74e-3, rel_tol=2e-3)
    assert p.robin_top == 1 and p.h_top == 10.0 and p.n_ext_zhi == 1.0
    assert np.all(prob.temp == 310.15)


def test_batch_manifest_is_the_committed_one():
    """The 64-instance manifest is drawn from a fixed PCG64 seed; the JSON
    beside the module is that draw (edges, optics and packet counts in the
    configured ranges)."""
    import json
    m = configs.batch_manifest()
    path = os.path.join(os.path.dirname(configs.__file__), "batch_manifest.json")
    assert json.load(open(path)) == m
    assert len(m) == 64 and {x["edge"] for x in m} <= {64, 96, 128, 160}
    assert all(0.1 <= x["mua"] <= 10 and 50 <= x["mus"] <= 300 and 0.7 <= x["g"] <= 0.95 for x in m)
    assert all(1e6 <= x["packets"] <= 1e7 for x in m)


def test_short_pulse_step_guard():
    """heat_sim_3D's guard (3dFD.f90:102-105): a pulse shorter than the
    explicit step shortens the step to pulse/100; the run length is set
    before it, from the original step (mcpolar.f90:133-140)."""
    from tam_amd import driver
    inp = dict(configs.REFERENCE_INPUT, energyPerPixel=0.01, pulsetype="tophat")
    prob = configs.reference_default(inp, 24)
    delt0, pl = prob.params.delt, prob.pulse_length
    assert math.isclose(pl, 0.01e-3 * 81 / 70.0, rel_tol=1e-15) and pl < delt0
    assert driver.step_guard(pl, delt0) == pl / 100.0
    assert driver.step_guard(pl, 0.5 * pl) == 0.5 * pl        # a longer pulse keeps the step
    assert driver.step_guard(pl, pl) == pl                    # strict comparison
    # the run length uses the original step: tophat, total_time 2 s is kept
    real_pulse, total = driver.run_length(inp, pl, delt0)
    assert real_pulse == pl and total == inp["total_time"]
    # the reference default (a 0.46 s pulse) is not affected
    ref = configs.reference_default(n=24)
    assert driver.step_guard(ref.pulse_length, ref.params.delt) == ref.params.delt

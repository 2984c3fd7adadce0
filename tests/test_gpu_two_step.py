"""Two heat steps per launch (temporal blocking) against one step per launch.

The two-step shapes (tiles 6-8) compute both steps with the one-step
arithmetic in the same order, so temperatures, Omega and threshold times are
bit-identical to the one-step shapes; only the boundary-energy sum is
accumulated in another order.  Cases:
* temperatures across the Arrhenius window with a removed crater, a source,
  Robin top and in-step damage: the threshold chain fires during the run, one
  threshold per voxel and step;
* a pulse train whose laser switches between the two steps of a launch;
* odd step counts (the last step runs alone) and kz from 1 to past nz.
"""
import numpy as np
import pytest

from tam_amd import configs, lib

pytestmark = pytest.mark.gpu

ONE_STEP = 4                # direct-32x8
TWO_STEP = (6, 7, 8)        # two-step-32x8, two-step-32x16, two-step-64x8


def _print(*a):
    print(*a, flush=True)


def _window_block(n=(45, 38, 27), seed=5):
    nx, ny, nz = n
    p = configs._config_params(n, (0.01 * nx, 0.01 * ny, 0.01 * nz))
    p.damage_every_step = 1
    p.arr_A, p.arr_dE = 3.1e98, 6.28e5
    mat = np.zeros((nz, ny, nx), np.uint8)
    temp, kappa, density, heatcap = configs._thermal_cube(p, configs.C0_THERMAL)
    rng = np.random.default_rng(seed)
    temp = temp.copy()
    temp[1:-1, 1:-1, 1:-1] = 325.0 + 30.0 * rng.random((nz, ny, nx))    # across the damage window
    rhokap = np.zeros((nz + 2, ny + 2, nx + 2))
    rhokap[1:-1, 1:-1, 1:-1] = 100.5
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    crater = ((xx - nx // 2) ** 2 + (yy - ny // 2) ** 2 <= 20) & (zz >= nz - 3)
    rhokap[1:-1, 1:-1, 1:-1][crater] = 0.0                                # removed: no conduction
    prob = configs.Problem("window", p, mat, [(0.5, 100.0, 0.9, 1.38)], temp, kappa, density, heatcap,
                           rhokap=rhokap)
    jm = 5e7 * rng.random((nz, ny, nx))
    jm[crater] = 0.0
    return prob, jm


def _pulse(delt):
    q = lib.Pulse()
    q.realPulseLength = 5 * delt
    q.repetitionRate_1 = 4 * delt
    q.pulsesToDo = 4
    q.laser_flag = 1
    q.laserOn = 1.0
    q.use_pulse = 1
    return q


def _run(prob, jm, tile, kz, steps, pulse=False):
    g = prob.grid()
    g.upload(lib.FIELD_JMEAN, jm)
    g.set_heat_launch(tile, kz)
    q = _pulse(prob.params.delt) if pulse else None
    ons = []
    if pulse:
        for _ in range(steps):                  # one step per call: the laser state of every step
            ons.append(q.laserOn)
            g.heat_run(1, q)
        g2 = prob.grid()                        # the same steps in one call (pairs + odd step)
        g2.upload(lib.FIELD_JMEAN, jm)
        g2.set_heat_launch(tile, kz)
        q2 = _pulse(prob.params.delt)
        g2.heat_run(steps, q2)
        assert (q2.time, q2.laserOn, q2.pulsesDone) == (q.time, q.laserOn, q.pulsesDone)
        out2 = g2.download(lib.FIELD_TEMP)
        g2.close()
    else:
        g.heat_run(steps)
    out = dict(T=g.download(lib.FIELD_TEMP), om=g.download(lib.FIELD_TISSUE), th=g.download(lib.FIELD_THRESTIME),
               out=g.energy()[1], class_mode=g.timing().class_mode)
    g.close()
    if pulse:
        out["T_one_call"] = out2
        out["ons"] = ons
    return out


@pytest.mark.parametrize("steps", [40, 41])
def test_two_step_damage_bit_identical(steps):
    _print(f"case: 45x38x27 window block, crater, Robin top, in-step damage, {steps} steps")
    prob, jm = _window_block()
    ref = _run(prob, jm, ONE_STEP, 8, steps)
    assert ref["class_mode"] == 1
    th = ref["th"]
    fired = [(th[k] > 0).sum() for k in range(3)]
    _print("threshold voxels", fired, "max T", ref["T"].max())
    assert fired[0] > 1000 and fired[1] > 1000 and fired[0] < th[0].size
    # threshold times spread over the run (the chain fires in many different steps)
    assert len(np.unique(th[0][th[0] > 0])) > 5
    for tile in TWO_STEP:
        for kz in (1, 5, 8, 64):
            r = _run(prob, jm, tile, kz, steps)
            for k in ("T", "om", "th"):
                assert np.array_equal(r[k], ref[k]), (tile, kz, k)
            assert abs(r["out"] - ref["out"]) <= 1e-6 * abs(ref["out"]) + 1e-12, (tile, kz)


def test_two_step_pulse_switches_inside_a_pair():
    prob, jm = _window_block(n=(33, 31, 19), seed=9)
    steps = 27
    _print(f"case: 33x31x19 window block, pulse train, {steps} steps")
    ref = _run(prob, jm, ONE_STEP, 4, steps, pulse=True)
    ons = np.array(ref["ons"])
    switches = np.nonzero(np.diff(ons))[0] + 1          # first step with the new laser state
    _print("laser on per step", ons.astype(int).tolist())
    assert (switches % 2 == 1).any(), "a switch must fall between the two steps of a pair"
    assert np.array_equal(ref["T_one_call"], ref["T"])
    for tile in TWO_STEP:
        r = _run(prob, jm, tile, 4, steps, pulse=True)
        assert np.array_equal(r["T"], ref["T"]), tile
        assert np.array_equal(r["T_one_call"], ref["T"]), tile
        assert np.array_equal(r["om"], ref["om"]) and np.array_equal(r["th"], ref["th"]), tile


@pytest.mark.parametrize("n", [(1, 1, 1), (2, 3, 4), (31, 7, 2), (61, 13, 3)])
def test_two_step_small_and_ragged_grids(n):
    _print(f"case: two-step shapes on {n}, 9 steps")
    prob, jm = _window_block(n=n, seed=sum(n))
    ref = _run(prob, jm, ONE_STEP, 2, 9)
    for tile in TWO_STEP:
        for kz in (1, 2, 7):
            r = _run(prob, jm, tile, kz, 9)
            for k in ("T", "om", "th"):
                assert np.array_equal(r[k], ref[k]), (n, tile, kz, k)

"""The captured coupling cycle (tam_cycle, one HIP graph replay per cycle)
against the same cycle issued one C-ABI call at a time.

Both runs use the same Philox streams and launch shapes.  The only difference
allowed is the summation order of the FP64 tally atomics (last-bit changes of
jmean), so temperatures agree to 1e-6 relative (FP32 path) and 1e-12 (FP64
path), damage to 1e-6, and the ablation mask and transport counters exactly.
"""
import numpy as np
import pytest

from tam_amd import configs, lib
from tam_amd.cycle import CycleRunner

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _skin(steps):
    prob = configs.config1(n=(40, 40, 40), packets=150_000, steps=steps, cycles=4)
    prob.power = 40.0                 # enough to ablate within a few cycles at this size
    return prob


# (steps per cycle, stepwise launch shape, graph launch shape, transport
# blocks per CU); shape 8 runs two steps per launch, 4 one (None: the
# default shape); 1 block per CU gives the largest LDS tally box (160 KiB).
# On shape 8 a replay makes ceil(steps / 2) ping-pong swaps: 2, 3, 6 and 30
# steps give an odd count of launches from an even step count or the reverse,
# so the host's buffer parity after a replay must follow the launches.
@pytest.mark.parametrize("steps,shape_a,shape_b,bpc", [(60, None, None, None), (61, None, None, None),
                                                       (61, 4, 8, None), (60, 8, 4, 1), (2, 8, 8, None),
                                                       (3, 8, 8, None), (6, 4, 8, None), (30, 8, 8, None)])
def test_graph_cycle_matches_stepwise_fp32(steps, shape_a, shape_b, bpc):
    _print(f"case: 40^3 skin, {steps} steps per cycle, 4 cycles, graph vs stepwise, shapes {shape_a} {shape_b}, "
           f"transport blocks per CU {bpc}")
    prob = _skin(steps)
    a = CycleRunner(prob, seed=7, tune=False, graph=False)
    b = CycleRunner(prob, seed=7, tune=False, graph=True)
    if shape_a is not None:
        a.grid.set_heat_launch(shape_a, 8)
        b.grid.set_heat_launch(shape_b, 8)
    if bpc is not None:
        a.grid.set_mc_launch(bpc, 32)
        b.grid.set_mc_launch(bpc, 32)
    for cyc in range(4):
        sa = a.run_cycle()
        sb = b.run_cycle()
        for k in ("absorbed", "specular", "missed", "crossings", "interactions", "tally_adds"):
            assert getattr(sa, k) == getattr(sb, k), (cyc, k)
        assert sb.packets == prob.packets
        Ta, Tb = a.grid.download(lib.FIELD_TEMP), b.grid.download(lib.FIELD_TEMP)
        np.testing.assert_allclose(Tb, Ta, rtol=1e-6, atol=0)
        np.testing.assert_allclose(b.grid.download(lib.FIELD_TISSUE), a.grid.download(lib.FIELD_TISSUE),
                                   rtol=1e-6, atol=1e-30)
        ma, mb = a.grid.download(lib.FIELD_MAT), b.grid.download(lib.FIELD_MAT)
        assert np.array_equal(ma, mb), cyc
        _print(f"cycle {cyc}: max T {Ta.max():.2f} K, removed {int((ma == len(prob.table)).sum())}")
    if steps >= 60:
        assert (ma == len(prob.table)).any(), "the case should ablate"
    a.close()
    b.close()


def test_graph_cycle_laser_off_keeps_jmean():
    """npackets = 0: no transport and no normalisation in that replay, so the
    heat source keeps the last cycle's jmean (the reference's jmeanGLOBAL)."""
    prob = _skin(30)
    a = CycleRunner(prob, seed=3, tune=False, graph=False)
    b = CycleRunner(prob, seed=3, tune=False, graph=True)
    a.run_cycle()
    b.run_cycle()
    j0 = b.grid.download(lib.FIELD_JMEAN)
    # stepwise: heat + rebuild only; graph: a replay with zero packets
    a.grid.heat_run(prob.steps)
    a.grid.rebuild()
    st = b.grid.cycle(0, 3, 99, prob.scale, prob.steps)
    assert st.packets == 0
    assert np.array_equal(b.grid.download(lib.FIELD_JMEAN), j0)
    np.testing.assert_allclose(b.grid.download(lib.FIELD_TEMP), a.grid.download(lib.FIELD_TEMP), rtol=1e-6)
    a.close()
    b.close()


def test_graph_cycle_reference_mode_with_pulse():
    """Reference physics (FP64 stencil, boiling, water model, Arrhenius at the
    reference cadence, top-hat pulse train) over 30 one-step cycles."""
    prob = configs.reference_default(n=24)
    p = prob.params
    _print("case: reference default 24^3, 30 cycles of 1 step, pulse train")

    def pulse():
        q = lib.Pulse()
        q.realPulseLength = 12 * p.delt
        q.repetitionRate_1 = 20 * p.delt
        q.pulsesToDo = 2
        q.laser_flag = 1
        q.laserOn = 1.0
        q.use_pulse = 1
        return q

    ga, gb = prob.grid(), prob.grid()
    pa, pb = pulse(), pulse()
    # 3e-3 of the input.params power (70 W into a 0.06 x 0.06 x 0.12 cm grid
    # vaporises the top voxels in one step): the top voxels reach the boiling
    # plateau (373 K) within the pulses without being ablated
    npk = 20_000
    scale = 3e-3 * prob.power / (npk * p.volume_voxel)
    for cyc in range(30):
        on = pa.laser_flag != 0
        assert on == (pb.laser_flag != 0)
        if on:
            ga.mc_run(npk, seed=5, stream=cyc)
            ga.normalise(scale)
        ga.heat_run(1, pa)
        ga.arrhenius(p.delt, pa.time)
        ga.rebuild()
        gb.cycle(npk if on else 0, 5, cyc, scale, 1, arr_delt=p.delt, pulse=pb, stats=(cyc % 7 == 6))
        assert (pa.time, pa.laserOn, pa.pulsesDone) == (pb.time, pb.laserOn, pb.pulsesDone)
    gb.synchronize()
    Ta, Tb = ga.download(lib.FIELD_TEMP), gb.download(lib.FIELD_TEMP)
    _print("max T", Ta.max(), "max |dT|", np.abs(Ta - Tb).max())
    np.testing.assert_allclose(Tb, Ta, rtol=1e-12)
    np.testing.assert_allclose(gb.download(lib.FIELD_TISSUE), ga.download(lib.FIELD_TISSUE), rtol=1e-9,
                               atol=1e-300)
    assert np.array_equal(gb.download(lib.FIELD_THRESTIME), ga.download(lib.FIELD_THRESTIME))
    assert Ta.max() > 300.0
    ga.close()
    gb.close()

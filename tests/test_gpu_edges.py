"""GPU edge cases: degenerate and odd grid shapes, and calls with no work.

* Heat on grids down to 1x1x1 (tiles larger than the grid, a single plane,
  lanes past the grid edge): every launch shape gives the same field, within
  1e-5 of the FP64 oracle.
* Transport on 1x1x1 and flat grids follows the oracle's packet fates on
  the same Philox streams.
* Zero packets / zero steps leave the state untouched.
"""
import numpy as np
import pytest

from tam_amd import configs, lib
from test_gpu_parity import _fractions_oracle, _oracle_for
from test_gpu_stencil import _oracle

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (2, 3, 4), (7, 1, 5), (33, 9, 2), (65, 2, 3)]


def _print(*a):
    print(*a, flush=True)


def _block(n, seed=0):
    nx, ny, nz = n
    p = configs._config_params(n, (0.01 * nx, 0.01 * ny, 0.01 * nz))
    mat = np.zeros((nz, ny, nx), np.uint8)
    temp, kappa, density, heatcap = configs._thermal_cube(p, configs.C0_THERMAL)
    rng = np.random.default_rng(seed)
    temp = temp.copy()
    temp[1:-1, 1:-1, 1:-1] += 5.0 * rng.random((nz, ny, nx))       # a non-uniform start
    rhokap = np.zeros((nz + 2, ny + 2, nx + 2))
    rhokap[1:-1, 1:-1, 1:-1] = 100.5                               # mu_t of the material, no removed voxels
    prob = configs.Problem("block", p, mat, [(0.5, 100.0, 0.9, 1.38)], temp, kappa, density, heatcap,
                           rhokap=rhokap)
    jm = 2e6 * rng.random((nz, ny, nx))
    return prob, jm


@pytest.mark.parametrize("n", SHAPES)
def test_heat_odd_grids_all_shapes(n):
    _print(f"case: heat on {n}, 30 steps, every launch shape")
    prob, jm = _block(n, seed=sum(n))
    ref = None
    for tile in range(9):
        for kz in (1, 2, 5):
            g = prob.grid()
            g.upload(lib.FIELD_JMEAN, jm)
            g.set_heat_launch(tile, kz)
            g.heat_run(30)
            T = g.download(lib.FIELD_TEMP)
            g.close()
            if ref is None:
                ref = T
            assert np.array_equal(T, ref), (n, tile, kz)
    To = _oracle(prob, jm, 30)
    inner = (slice(1, -1),) * 3
    rel = np.abs(ref[inner] - To[inner]) / To[inner]
    _print("max rel vs oracle", rel.max())
    assert rel.max() <= 1e-5


@pytest.mark.parametrize("n", [(1, 1, 1), (9, 7, 1), (4, 4, 2)])
def test_mc_degenerate_grids_follow_oracle(n):
    nx, ny, nz = n
    p = configs._config_params(n, (0.02 * nx, 0.02 * ny, 0.02 * nz))
    p.beam = lib.BEAM_TOPHAT
    p.beam_radius = 0.5 * 0.02 * min(nx, ny)
    p.n_ext_zhi = 1.0
    mat = np.zeros((nz, ny, nx), np.uint8)
    temp, kappa, density, heatcap = configs._thermal_cube(p, configs.C0_THERMAL)
    prob = configs.Problem("tiny", p, mat, [(3.0, 90.0, 0.8, 1.4)], temp, kappa, density, heatcap,
                           packets=100_000)
    _print(f"case: transport on {n}, 1e5 packets, same stream as the oracle")
    g = prob.grid()
    st = g.mc_run(prob.packets, seed=99)
    t_gpu = g.download(lib.FIELD_TALLY)
    g.close()
    t_or, ost = _oracle_for(prob, 99).run(prob.packets)
    fg, fo = st.fractions(), _fractions_oracle(ost, prob.packets)
    _print("gpu", fg, "\noracle", fo)
    for k in fg:
        assert abs(fg[k] - fo[k]) <= 2e-3, k
    assert st.packets == st.absorbed + st.specular + sum(st.escaped) + st.missed + st.killed
    if t_or.sum() > 0:
        assert np.abs(t_gpu - t_or).sum() / t_or.sum() < 2e-2


def test_zero_work_leaves_state():
    prob, jm = _block((12, 10, 8), seed=3)
    g = prob.grid()
    g.upload(lib.FIELD_JMEAN, jm)
    T0 = g.download(lib.FIELD_TEMP)
    st = g.mc_run(0, seed=1)
    assert st.packets == 0 and st.absorbed == 0 and st.crossings == 0
    assert not g.download(lib.FIELD_TALLY).any()
    g.heat_run(0)
    assert np.array_equal(g.download(lib.FIELD_TEMP), T0)
    g.synchronize()
    g.close()

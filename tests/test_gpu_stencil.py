"""GPU tests of the FP32 stencil's two coefficient forms and launch shapes.

* Every (tile, kz) launch shape computes each voxel with the same arithmetic,
  so all shapes give bit-identical fields.
* Class mode (uint8 classes + LDS face table) and array mode (FP32 face
  arrays) hold bit-identical coefficients; both stay within 1e-5 of the FP64
  oracle, including a crater of removed (thermally inert) voxels.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi as O
from tam_amd import configs, lib

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _crater_problem(nx=37, ny=29, nz=23, random_props=False, seed=0):
    """Config-style block (Robin top, Dirichlet elsewhere) with a removed crater."""
    p = configs._config_params((nx, ny, nz), (0.37, 0.29, 0.23))
    mat = np.zeros((nz, ny, nx), np.uint8)
    table = [(0.5, 100.0, 0.9, 1.38)]
    temp, kappa, density, heatcap = configs._thermal_cube(p, configs.C0_THERMAL)
    rng = np.random.default_rng(seed)
    if random_props:
        f = lambda v: v * (1 + 0.2 * (rng.random(v.shape) - 0.5))
        kappa, density, heatcap = f(kappa), f(density), f(heatcap)
    rhokap = np.zeros((nz + 2, ny + 2, nx + 2))
    rhokap[1:-1, 1:-1, 1:-1] = 100.5
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    crater = ((xx - nx // 2) ** 2 + (yy - ny // 2) ** 2 <= 16) & (zz >= nz - 4)
    rhokap[1:-1, 1:-1, 1:-1][crater] = 0.0
    temp = temp.copy()
    temp[1:-1, 1:-1, 1:-1][crater] = 400.0                 # removed voxels keep their T
    prob = configs.Problem("crater", p, mat, table, temp, kappa, density, heatcap, rhokap=rhokap)
    jmean = 3e6 * np.exp(-((xx - nx // 2) ** 2 + (yy - ny // 2) ** 2) / 30.0 - (nz - 1 - zz) / 5.0)
    jmean[crater] = 0.0
    return prob, jmean


def _gpu_run(prob, jmean, steps, tile=None, kz=None):
    g = prob.grid()
    g.upload(lib.FIELD_JMEAN, jmean)
    if tile is not None:
        g.set_heat_launch(tile, kz)
    g.heat_run(steps)
    T = g.download(lib.FIELD_TEMP)
    t = g.timing()
    g.close()
    return T, t


def _oracle(prob, jmean, steps):
    p = prob.params
    coeff = np.zeros_like(prob.temp)
    coeff[1:-1, 1:-1, 1:-1] = p.delt / (prob.density * prob.heatcap)[1:-1, 1:-1, 1:-1]
    T = np.ascontiguousarray(prob.temp.copy())
    arrs = [np.ascontiguousarray(a) for a in (prob.kappa, prob.density, prob.heatcap, coeff, jmean,
                                              prob.rhokap)]
    Q = np.zeros(jmean.shape)
    h = O.Heat()
    h.nx, h.ny, h.nz = p.nx, p.ny, p.nz
    h.dx, h.dy, h.dz, h.delt, h.dscale = p.dx_m, p.dy_m, p.dz_m, p.delt, p.dscale
    h.temp = T.ctypes.data
    h.kappa, h.density, h.heatcap, h.coeff, h.jmean, h.rhokap = [a.ctypes.data for a in arrs]
    h.Q = Q.ctypes.data
    h.volumeVoxel, h.massVoxel, h.QVapor = p.volume_voxel, p.mass_voxel, p.qvapor
    h.robin_top, h.h_top, h.T_amb = p.robin_top, p.h_top, p.t_amb
    h.removed_inert = p.removed_inert
    assert O.lib().or_heat_steps(C.byref(h), steps, None) == 0
    return T


@pytest.mark.parametrize("random_props", [False, True])
def test_launch_shapes_bit_identical(random_props):
    prob, jm = _crater_problem(random_props=random_props)
    for steps in (40, 41):                      # 41: a two-step run ends with one step alone
        ref, t0 = _gpu_run(prob, jm, steps, tile=0, kz=4)
        _print("class mode" if t0.class_mode else "array mode", steps, "steps")
        assert bool(t0.class_mode) == (not random_props)
        # 0-2: LDS planes, 3-5: direct neighbour loads, 6-8: two steps per launch
        for tile in range(9):
            for kz in (1, 3, 8, 32):
                T, _ = _gpu_run(prob, jm, steps, tile=tile, kz=kz)
                assert np.array_equal(T, ref), (steps, tile, kz)


@pytest.mark.parametrize("random_props", [False, True])
def test_crater_heat_within_1e5(random_props):
    prob, jm = _crater_problem(random_props=random_props, seed=3)
    Tg, t = _gpu_run(prob, jm, 500)
    To = _oracle(prob, jm, 500)
    inner = (slice(1, -1),) * 3
    rel = np.abs(Tg[inner] - To[inner]) / To[inner]
    _print("mode", "class" if t.class_mode else "array", "max rel", rel.max(),
           "max rise", (To[inner] - prob.params.t_ref).max())
    assert rel.max() <= 1e-5
    crater = prob.rhokap[1:-1, 1:-1, 1:-1] == 0
    # inert voxels keep their temperature (as held on the device: t_ref + FP32(T - t_ref))
    t_ref = prob.params.t_ref
    kept = t_ref + (prob.temp[inner][crater] - t_ref).astype(np.float32).astype(np.float64)
    assert np.array_equal(Tg[inner][crater], kept)


def test_class_and_array_modes_agree():
    """Same physics through both coefficient forms: the class table is
    exercised by uniform properties; array mode by the same properties with a
    harmless 16th distinct value in one halo corner cell (never read)."""
    prob, jm = _crater_problem()
    Tc, tc = _gpu_run(prob, jm, 60)
    assert tc.class_mode == 1
    hs = prob.kappa.shape
    kappa = prob.kappa.copy()
    # 15 extra distinct conductivities in unused halo corner cells force array mode
    corners = [(0, 0, 0), (0, 0, hs[2] - 1), (0, hs[1] - 1, 0), (hs[0] - 1, 0, 0),
               (0, hs[1] - 1, hs[2] - 1), (hs[0] - 1, hs[1] - 1, 0), (hs[0] - 1, 0, hs[2] - 1),
               (hs[0] - 1, hs[1] - 1, hs[2] - 1)]
    for n, cidx in enumerate(corners):
        kappa[cidx] = 0.3 + 0.01 * n
    for n in range(8):
        kappa[0, 0, 1 + n] = 0.2 + 0.01 * n          # z=0,y=0 edge: never a face of an interior voxel
    prob2 = configs.Problem("crater-array", prob.params, prob.mat, prob.table, prob.temp, kappa,
                            prob.density, prob.heatcap, rhokap=prob.rhokap)
    Ta, ta = _gpu_run(prob2, jm, 60)
    assert ta.class_mode == 0
    assert np.array_equal(Tc, Ta)


@pytest.mark.parametrize("damage", [0, 1])
def test_autotune_keeps_state(damage):
    prob, jm = _crater_problem()
    prob.params.damage_every_step = damage
    g = prob.grid()
    g.upload(lib.FIELD_JMEAN, jm)
    if damage:                                  # a damage state with thresholds part crossed
        T = prob.temp.copy()
        T[1:-1, 1:-1, 1:-1] = 335.0 + 20.0 * np.random.default_rng(4).random(T[1:-1, 1:-1, 1:-1].shape)
        g.upload(lib.FIELD_TEMP, T)
        g.heat_run(5)
    fields = (lib.FIELD_TEMP, lib.FIELD_TISSUE, lib.FIELD_THRESTIME, lib.FIELD_Q)
    before = [g.download(f) for f in fields]
    if damage:
        assert (before[2] > 0).any() and (before[1] > 0).any()
    tile, kz, ms = g.tune_heat(reps=3)
    after = [g.download(f) for f in fields]
    g.heat_run(3)                               # the live state still steps after tuning
    g.close()
    _print("tuned", lib.__name__, tile, kz, ms)
    assert 3 <= tile < 9 and kz >= 1 and ms > 0  # class mode: direct and two-step shapes
    for f, b, a in zip(fields, before, after):
        assert np.array_equal(b, a), f

"""The HIP path against closed forms and the normalisation formula, through
the C-ABI (independent of the oracle):

  * normalisation (mcpolar.f90:173-174): jmean = tally * mu_a * scale per
    voxel, bit for bit against numpy on the downloaded tally; and the
    track-length identity E[sum_v mu_a L_v] = absorbed packets, which ties the
    tally to the fates (within 4 standard errors over independent streams);
  * the FP32 stencil's convective top face: the steady profile of a column
    over a Dirichlet plane (slope fixed by flux continuity at the face);
  * the FP32 stencil's half-node means: the series-resistance steady state of
    a column of random kappa, rho and c (class-table and per-voxel modes);
  * the transport's Fresnel reflection: a pure absorber between two
    mismatched faces under normal incidence (multiple-reflection series).

Tolerances are stated in each test.
"""
import math

import numpy as np
import pytest

from tam_amd import configs, lib

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


# --------------------------------------------------------- normalisation --

def _crater_skin(n=40, packets=200_000):
    prob = configs.three_layer_skin(n=(n, n, n), ext=(1.0, 1.0, 0.5), packets=packets, steps=1, cycles=1)
    mut = np.array([t[0] + t[1] for t in prob.table])
    rk = np.zeros((n + 2,) * 3)
    rk[1:-1, 1:-1, 1:-1] = mut[prob.mat]
    zz, yy, xx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    crater = ((xx - n / 2 + 0.5) ** 2 + (yy - n / 2 + 0.5) ** 2 <= 25) & (zz >= n - 6)
    rk[1:-1, 1:-1, 1:-1][crater] = 0.0                 # ablated: mu_t = 0, the material id kept
    prob.rhokap = rk
    return prob


def test_normalise_is_tally_times_mua_times_scale():
    """jmean_v = L_v * (rhokap_v * mu_a/mu_t of its material) * scale,
    computed in that order, bit-identical to numpy; 0 in ablated voxels."""
    prob = _crater_skin()
    _print("case: 40^3 skin with crater, 2e5 packets, normalise vs numpy")
    g = prob.grid()
    g.mc_run(prob.packets, seed=99)
    tally = g.download(lib.FIELD_TALLY)
    scale = 3.7 / (prob.packets * prob.params.volume_voxel)
    g.normalise(scale)
    jm = g.download(lib.FIELD_JMEAN)
    rk = g.download(lib.FIELD_RHOKAP)[1:-1, 1:-1, 1:-1]
    mat = g.download(lib.FIELD_MAT).astype(np.int64)
    g.close()
    frac = np.array([t[0] / (t[0] + t[1]) for t in prob.table] + [0.0])
    expect = tally * (rk * frac[mat]) * scale
    assert np.array_equal(jm, expect)
    assert (rk == 0).any() and not jm[rk == 0].any() and jm.max() > 0


def test_track_length_estimator_matches_absorbed_packets():
    """Summed over voxels, mu_a * path length is an unbiased estimate of the
    absorbed packets: over 10 independent streams the mean difference per
    packet is within 4 standard errors of zero (and the power it normalises
    to, sum_v jmean_v V = P * A, with it)."""
    prob = _crater_skin(packets=100_000)
    g = prob.grid()
    frac = np.array([t[0] / (t[0] + t[1]) for t in prob.table] + [0.0])
    d, A = [], []
    scale = 1.0 / (prob.packets * prob.params.volume_voxel)          # P = 1 W
    for b in range(10):
        st = g.mc_run(prob.packets, seed=5, stream=b)
        g.normalise(scale)
        jm = g.download(lib.FIELD_JMEAN)
        mat = g.download(lib.FIELD_MAT).astype(np.int64)
        rk = g.download(lib.FIELD_RHOKAP)[1:-1, 1:-1, 1:-1]
        est = (g.download(lib.FIELD_TALLY) * rk * frac[mat]).sum() / prob.packets
        assert math.isclose(jm.sum() * prob.params.volume_voxel, est, rel_tol=1e-12)
        d.append(est - st.absorbed / prob.packets)
        A.append(st.absorbed / prob.packets)
    g.close()
    d = np.array(d)
    se = d.std(ddof=1) / math.sqrt(len(d))
    _print("absorbed", np.mean(A), "mean diff", d.mean(), "se", se)
    assert abs(d.mean()) < 4 * se + 1e-12
    assert np.mean(A) > 0.3


# ------------------------------------------------------- stencil closed forms --

def _column_params(nxy, nz, dz, delt, dscale, robin=None):
    p = configs.base_params(nxy, nxy, nz, nxy * dz * 50, nxy * dz * 50, nz * dz * 50)
    p.dx_m = p.dy_m = p.dz_m = dz
    p.delt, p.dscale = delt, dscale
    p.volume_voxel = dz ** 3
    p.mass_voxel = 1000.0 * p.volume_voxel
    p.qvapor = 1e30
    p.boil = 0
    p.removed_inert = 1                  # (no voxel is removed) enables the class-table mode
    p.damage_every_step = 0
    p.thermal_mode = lib.THERMAL_TABLE
    if robin is not None:
        p.robin_top, p.h_top, p.t_amb = 1, robin[0], robin[1]
    return p


def _column(nz, profile, nxy):
    return np.broadcast_to(np.asarray(profile, float)[:, None, None], (nz + 2, nxy + 2, nxy + 2)).copy()


def _run_column(p, T, K, D, H, steps, tile=None):
    g = lib.TissueGrid(p)
    mat = np.zeros((p.nz, p.ny, p.nx), np.uint8)
    g.set_optics(mat, [(0.5, 100.0, 0.9, 1.38)])
    g.set_thermal(T, K, D, H)
    if tile is not None:
        g.set_heat_launch(tile, 4)
    g.heat_run(steps)
    out = g.download(lib.FIELD_TEMP)
    mode = g.timing().class_mode
    g.close()
    return out, mode


@pytest.mark.parametrize("tile", [None, 4, 8])
def test_robin_top_steady_state_closed_form(tile):
    """Column of 12 planes over a Dirichlet plane T0 = 310 K, convective top
    face (h = 5000 W/m^2/K, T_amb = 298 K, h dz / kappa = 1): the field
    settles to T_k = T0 + s k, s = h (T_amb - T0) / (kappa/dz + h (N + 1/2)).
    FP32 increments: within 1e-4 of |T0 - T_amb| (the h -> 2h/2 mistake
    moves the top plane by 0.38 K)."""
    N, nxy, dz = 12, 3, 1e-4
    kappa, rho, cp = 0.5, 1000.0, 4000.0
    h, T0, Tamb = 5000.0, 310.0, 298.0
    s = h * (Tamb - T0) / (kappa / dz + h * (N + 0.5))
    exact = T0 + s * np.arange(N + 2)
    dscale = 1.0
    delt = 0.9 / (2 * dscale * kappa / (rho * cp) * 3 / dz ** 2)
    p = _column_params(nxy, N, dz, delt, dscale, robin=(h, Tamb))
    p.t_ref = T0
    T = _column(N, exact, nxy)
    T[1:-1, 1:-1, 1:-1] = T0
    ones = np.ones_like(T)
    _print(f"case: Robin column 3x3x12, 6000 steps, tile {tile}")
    Tn, _ = _run_column(p, T, kappa * ones, rho * ones, cp * ones, 6000, tile)
    err = np.abs(Tn[1:-1, 1:-1, 1:-1] - exact[1:-1, None, None]).max()
    _print("max |T - closed form|", err)
    assert err < 1e-4 * abs(T0 - Tamb)


@pytest.mark.parametrize("N", [12, 20])
def test_series_resistance_steady_state_random_properties(N):
    """Random kappa, rho and c per plane (halo planes included) between two
    Dirichlet planes: the steady profile follows the series resistances
    1 / (kbar / (rhobar cbar)) of the faces.  12 planes fit the 16-class
    table, 20 run on per-voxel coefficients.  FP32: within 1e-4 of the span."""
    rng = np.random.default_rng(N)
    nxy, dz = 3, 1e-4
    kap = 0.5 * (0.3 + 1.4 * rng.random(N + 2))
    rho = 1000.0 * (0.5 + rng.random(N + 2))
    cp = 4000.0 * (0.5 + rng.random(N + 2))
    Tb, Tt = 300.0, 320.0
    m = lambda a: 0.5 * (a[:-1] + a[1:])
    G = m(kap) / (m(rho) * m(cp))
    R = np.concatenate([[0.0], np.cumsum(1.0 / G)])
    exact = Tb + (Tt - Tb) * R / R[-1]
    dscale = 0.5
    a = dscale * G / dz ** 2
    lim = 1.0 / ((a[:-1] + a[1:]).max() + 4 * dscale * (kap / (rho * cp)).max() / dz ** 2)
    p = _column_params(nxy, N, dz, 0.9 * lim, dscale)
    p.t_ref = Tb
    T = _column(N, exact, nxy)
    T[1:-1, 1:-1, 1:-1] = Tb
    K, D, H = (_column(N, x, nxy) for x in (kap, rho, cp))
    Tn, mode = _run_column(p, T, K, D, H, 12000 if N == 12 else 30000)
    err = np.abs(Tn[1:-1, 1:-1, 1:-1] - exact[1:-1, None, None]).max()
    _print(f"case: series column N={N}, class mode {mode}, max |T - closed form| {err}")
    assert err < 1e-4 * (Tt - Tb)


# ---------------------------------------------------------- Fresnel slab --

@pytest.mark.parametrize("n_in,n_top,n_bot", [(2.4, 1.0, 1.2), (1.4, 0.0, 3.0)])
def test_fresnel_slab_multiple_reflections_closed_form(n_in, n_top, n_bot):
    """Normal incidence on a pure absorber (mu_a 2 /cm, 0.5 cm) between
    faces of reflectance ((n_i - n_t)/(n_i + n_t))^2: fates within 4 sigma
    (binomial) and the path length per plane within 4.5 standard errors over
    8 independent streams, against the geometric series of the reflections."""
    nz, zmax, mua = 20, 0.25, 2.0
    L, dz = 2 * zmax, 2 * zmax / nz
    R = lambda a, b: 0.0 if b <= 0 else ((a - b) / (a + b)) ** 2
    Rt, Rb = R(n_in, n_top), R(n_in, n_bot)
    t = math.exp(-mua * L)
    D = 1 - Rb * Rt * t * t
    enter = 1 - Rt
    fates = dict(specular=Rt, bottom=enter * (1 - Rb) * t / D, top=enter * Rb * t * t * (1 - Rt) / D)
    fates["absorbed"] = 1 - sum(fates.values())
    k = np.arange(nz)
    z1 = (nz - 1 - k) * dz
    z2 = z1 + dz
    plane = enter * ((np.exp(-mua * z1) - np.exp(-mua * z2)) +
                     Rb * t * (np.exp(-mua * (L - z2)) - np.exp(-mua * (L - z1)))) / (mua * D)
    p = configs._config_params((5, 5, nz), (1.0, 1.0, 2 * zmax))
    p.beam, p.beam_radius = lib.BEAM_TOPHAT, 0.05
    p.n_ext_zhi, p.n_ext_zlo = n_top, n_bot
    mat = np.zeros((nz, 5, 5), np.uint8)
    temp, kappa, density, heatcap = configs._thermal_cube(p, configs.C0_THERMAL)
    prob = configs.Problem("slab", p, mat, [(mua, 0.0, 0.0, n_in)], temp, kappa, density, heatcap)
    g = prob.grid()
    nb, per = 8, 250_000
    planes, acc = [], dict(specular=0.0, bottom=0.0, top=0.0, absorbed=0.0)
    for b in range(nb):
        st = g.mc_run(per, seed=31, stream=b)
        planes.append(g.download(lib.FIELD_TALLY).sum(axis=(1, 2)) / per)
        acc["specular"] += st.specular
        acc["bottom"] += st.escaped[lib.FACE_ZLO]
        acc["top"] += st.escaped[lib.FACE_ZHI]
        acc["absorbed"] += st.absorbed
        assert st.missed == 0 and sum(st.escaped[:4]) == 0 and st.killed == 0
    g.close()
    N = nb * per
    for name, pr in fates.items():
        sig = math.sqrt(max(pr * (1 - pr), 1e-12) / N)
        _print(name, acc[name] / N, pr, sig)
        assert abs(acc[name] / N - pr) < 4 * sig + 1e-12, name
    planes = np.array(planes)
    se = planes.std(axis=0, ddof=1) / math.sqrt(nb)
    z = np.abs(planes.mean(axis=0) - plane) / se
    _print("max z per plane", z.max())
    assert z.max() < 4.5

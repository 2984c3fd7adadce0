"""Closed-form pins of the oracle branches that every configuration runs
(CPU only): the convective (Robin) top face, the half-node property means of
a medium whose conductivity, density and heat capacity vary from voxel to
voxel, and the Fresnel internal reflection of the photon walk.

Each check is chosen so that a plausible restatement mistake fails it:
  * Robin ghost node (3dFD.f90 has none; the extension of the configurations):
    the steady profile of a column with a Dirichlet bottom is linear with the
    slope fixed by flux continuity at the face half a node above the last
    plane, kappa s / dz = h (T_amb - T_face).  h in place of h/2, T_N in place
    of the face mean or a flipped sign all move the fixed point.
  * half-node means (3dFD.f90:119-161): a column of random kappa, rho and c
    between two Dirichlet planes settles to the series-resistance profile of
    the face conductances kbar / (rhobar cbar) (flux continuity); a harmonic
    mean, a mean of the diffusivities or a swapped neighbour gives another
    profile.  The update is also reciprocal (a unit impulse at p raises q by
    exactly what one at q raises p), and with insulating faces it conserves
    heat content exactly.
  * Fresnel reflection in the walk (extension): a normally incident beam in a
    pure absorber between two mismatched faces reflects back and forth; the
    fates and the path length per plane follow the geometric series of the
    two reflectances.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_ffi as O


def _run_heat(T, kappa, rho, cp, delt, dscale, dx, dy, dz, steps, jmean=None, robin=None):
    """or_heat_steps on halo arrays indexed [z, y, x]; returns the new field."""
    nz, ny, nx = (s - 2 for s in T.shape)
    h = O.Heat()
    h.nx, h.ny, h.nz = nx, ny, nz
    h.dx, h.dy, h.dz = dx, dy, dz
    h.delt, h.dscale = delt, dscale
    coeff = np.zeros_like(T)
    coeff[1:-1, 1:-1, 1:-1] = delt / (rho * cp)[1:-1, 1:-1, 1:-1]
    arrs = [np.array(a, dtype=float, order="C") for a in (T, kappa, rho, cp, coeff)]
    jm = np.ascontiguousarray(np.zeros((nz, ny, nx)) if jmean is None else jmean, dtype=float)
    Q = np.zeros((nz, ny, nx))
    h.temp, h.kappa, h.density, h.heatcap, h.coeff = [a.ctypes.data for a in arrs]
    h.jmean, h.Q = jm.ctypes.data, Q.ctypes.data
    V = dx * dy * dz
    h.volumeVoxel, h.massVoxel, h.QVapor = V, float(rho.flat[0]) * V, 1e30
    h.boil = 0
    if robin is not None:
        h.robin_top, h.h_top, h.T_amb = 1, robin[0], robin[1]
    assert O.lib().or_heat_steps(C.byref(h), steps, None) == 0
    return arrs[0]


def _column(nz, profile, nxy=3):
    """Halo field whose every x-y position (halo included) carries `profile`
    (length nz + 2, bottom halo first): lateral terms vanish when the
    interior follows it."""
    return np.broadcast_to(np.asarray(profile, float)[:, None, None], (nz + 2, nxy + 2, nxy + 2)).copy()


# ------------------------------------------------------- Robin top face --

@pytest.mark.parametrize("dscale", [0.5, 1.0])
def test_robin_top_steady_state_closed_form(dscale):
    """Column of N planes over a Dirichlet plane T0, convective top face to
    T_amb with coefficient h.  The ghost node sits one node above plane N and
    the face half-way between them, so the steady profile is T_k = T0 + s k
    with kappa s / dz = h (T_amb - (T_N + s/2)):
        s = h (T_amb - T0) / (kappa/dz + h (N + 1/2)).
    The run starts from a uniform field and must settle there."""
    N, dz = 12, 1e-4
    kappa, rho, cp = 0.5, 1000.0, 4000.0
    h, T0, Tamb = 5000.0, 310.0, 298.0            # h dz / kappa = 1: the ghost term is not small
    s = h * (Tamb - T0) / (kappa / dz + h * (N + 0.5))
    k = np.arange(N + 2)
    exact = T0 + s * k
    T = _column(N, exact)
    T[1:-1, 1:-1, 1:-1] = T0                       # interior starts uniform
    ones = np.ones_like(T)
    alpha = kappa / (rho * cp)
    delt = 0.9 / (2 * dscale * alpha * 3 / dz ** 2)
    Tn = _run_heat(T, kappa * ones, rho * ones, cp * ones, delt, dscale, dz, dz, dz, 6000, robin=(h, Tamb))
    col = Tn[1:-1, 2, 2]
    assert np.abs(col - exact[1:-1]).max() < 1e-9 * abs(T0 - Tamb)
    # the same field is a fixed point in every interior column
    assert np.abs(Tn[1:-1, 1:-1, 1:-1] - exact[1:-1, None, None]).max() < 1e-9 * abs(T0 - Tamb)
    # h -> 0 would leave the column at T0; h/2 -> h would give another slope
    s_wrong = h * (Tamb - T0) / (kappa / dz + h * (N + 1.0))
    assert abs(col[-1] - (T0 + s_wrong * N)) > 1e-3


# ------------------------------------------- varying-coefficient means --

def _face_conductance(kappa, rho, cp):
    """kbar / (rhobar cbar) of each face between consecutive planes."""
    m = lambda a: 0.5 * (a[:-1] + a[1:])
    return m(kappa) / (m(rho) * m(cp))


def test_series_resistance_steady_state_random_properties():
    """Random kappa, rho and c per plane (halo planes included); Dirichlet
    planes below and above.  At steady state every face carries the same
    flux G_f (T_{k+1} - T_k), so T_k - T_0 is proportional to the sum of 1/G_f
    over the faces below plane k."""
    rng = np.random.default_rng(21)
    N, dz = 16, 1e-4
    kap = 0.5 * (0.3 + 1.4 * rng.random(N + 2))
    rho = 1000.0 * (0.5 + rng.random(N + 2))
    cp = 4000.0 * (0.5 + rng.random(N + 2))
    Tb, Tt = 300.0, 320.0
    G = _face_conductance(kap, rho, cp)
    R = np.concatenate([[0.0], np.cumsum(1.0 / G)])
    exact = Tb + (Tt - Tb) * R / R[-1]
    # harmonic-mean and mean-diffusivity variants give clearly other profiles
    hm = lambda a: 2 * a[:-1] * a[1:] / (a[:-1] + a[1:])
    m = lambda a: 0.5 * (a[:-1] + a[1:])
    for Gw in (hm(kap) / (m(rho) * m(cp)), m(kap / (rho * cp))):
        Rw = np.concatenate([[0.0], np.cumsum(1.0 / Gw)])
        assert np.abs(Tb + (Tt - Tb) * Rw / Rw[-1] - exact).max() > 1e-2
    T = _column(N, exact)
    T[1:-1, 1:-1, 1:-1] = Tb
    K, D, H = (_column(N, a) for a in (kap, rho, cp))
    dscale = 0.5
    # explicit limit of the stiffest plane: the sum of its six face coefficients
    a = dscale * G / dz ** 2
    lim = 1.0 / (np.maximum(a[:-1] + a[1:], 0).max() + 4 * dscale * (kap / (rho * cp)).max() / dz ** 2)
    Tn = _run_heat(T, K, D, H, 0.9 * lim, dscale, dz, dz, dz, 20000)
    assert np.abs(Tn[1:-1, 1:-1, 1:-1] - exact[1:-1, None, None]).max() < 1e-9 * (Tt - Tb)


def test_update_is_reciprocal_with_random_properties():
    """Half-node means are symmetric in the two cells, so the one-step
    response at q to a unit impulse at p equals the response at p to one at
    q, bit for bit, in each of the six directions (zero Dirichlet halo)."""
    rng = np.random.default_rng(5)
    n = 6
    hs = (n + 2,) * 3
    kap = 0.5 * (0.5 + rng.random(hs))
    rho = 1000.0 * (0.5 + rng.random(hs))
    cp = 4000.0 * (0.5 + rng.random(hs))
    dx, dy, dz = 1e-4, 1.3e-4, 0.8e-4
    p = (3, 3, 3)
    for off in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        q = tuple(a + b for a, b in zip(p, off))
        Ta = np.zeros(hs)
        Ta[p] = 1.0
        Tb = np.zeros(hs)
        Tb[q] = 1.0
        ra = _run_heat(Ta, kap, rho, cp, 1e-3, 0.5, dx, dy, dz, 1)
        rb = _run_heat(Tb, kap, rho, cp, 1e-3, 0.5, dx, dy, dz, 1)
        assert ra[q] > 0.0 and ra[q] == rb[p], off
        # a second-nearest cell gets nothing in one step
        assert ra[tuple(a + 2 * b for a, b in zip(p, off))] == 0.0


def test_insulated_random_conductivity_conserves_heat_content():
    """Random interior kappa, uniform rho c; every halo cell carries minus the
    conductivity of its interior neighbour, so each boundary face mean is
    exactly zero (an insulated box).  The fluxes between interior cells
    telescope, and the heat content grows by exactly the source energy."""
    rng = np.random.default_rng(8)
    n, dx = 10, 1e-4
    hs = (n + 2,) * 3
    kap = np.zeros(hs)
    kap[1:-1, 1:-1, 1:-1] = 0.5 * (0.2 + 1.6 * rng.random((n, n, n)))
    kap[0, 1:-1, 1:-1] = -kap[1, 1:-1, 1:-1]
    kap[-1, 1:-1, 1:-1] = -kap[-2, 1:-1, 1:-1]
    kap[1:-1, 0, 1:-1] = -kap[1:-1, 1, 1:-1]
    kap[1:-1, -1, 1:-1] = -kap[1:-1, -2, 1:-1]
    kap[1:-1, 1:-1, 0] = -kap[1:-1, 1:-1, 1]
    kap[1:-1, 1:-1, -1] = -kap[1:-1, 1:-1, -2]
    rho, cp = 1000.0, 4000.0
    T = 300.0 + 40.0 * rng.random(hs)
    jm = 1e8 * rng.random((n, n, n))
    delt, steps = 0.3 * dx ** 2 * rho * cp / (6 * kap.max()), 150
    Tn = _run_heat(T, kap, rho * np.ones(hs), cp * np.ones(hs), delt, 1.0, dx, dx, dx, steps, jmean=jm)
    V = dx ** 3
    dE = ((Tn - T)[1:-1, 1:-1, 1:-1] * rho * cp * V).sum()
    src = jm.sum() * V * delt * steps
    assert abs(dE - src) < 1e-11 * (rho * cp * V * T[1:-1, 1:-1, 1:-1].sum())
    assert np.abs(Tn - T).max() > 1.0                      # the field did move


# ------------------------------------------------- Fresnel in the walk --

def _slab_closed_form(mua, L, Rt, Rb, nz, dz):
    """Normal incidence on a pure absorber of thickness L between faces of
    normal-incidence reflectance Rt (top, also the launch specular) and Rb:
    fates per launched packet and the expected path length per plane
    (plane 0 at the bottom)."""
    t = math.exp(-mua * L)
    D = 1.0 - Rb * Rt * t * t
    enter = 1.0 - Rt
    fates = {"specular": Rt, "bottom": enter * (1 - Rb) * t / D, "top": enter * Rb * t * t * (1 - Rt) / D}
    fates["absorbed"] = 1.0 - sum(fates.values())
    k = np.arange(nz)
    z1 = (nz - 1 - k) * dz                          # depth of the plane's upper face
    z2 = z1 + dz
    down = np.exp(-mua * z1) - np.exp(-mua * z2)
    up = Rb * t * (np.exp(-mua * (L - z2)) - np.exp(-mua * (L - z1)))
    return fates, enter * (down + up) / (mua * D)


@pytest.mark.parametrize("n_in,n_top,n_bot", [(2.4, 1.0, 1.2), (1.4, 1.0, 1.0), (1.4, 0.0, 3.0)])
def test_fresnel_slab_multiple_reflections_closed_form(n_in, n_top, n_bot):
    """Every packet moves along z only; it reflects at each face with the
    normal-incidence Fresnel reflectance ((n_i - n_t) / (n_i + n_t))^2
    (n_ext <= 0: index matched, no reflection).  Fates within 4 sigma
    (binomial) and the path length per plane within 4 sigma (8 batches)."""
    nz, zmax, mua = 20, 0.25, 2.0
    L, dz = 2 * zmax, 2 * zmax / nz
    R = lambda a, b: 0.0 if b <= 0 else ((a - b) / (a + b)) ** 2
    Rt, Rb = R(n_in, n_top), R(n_in, n_bot)
    fates, plane = _slab_closed_form(mua, L, Rt, Rb, nz, dz)
    mat = np.zeros((nz, 5, 5), np.uint8)
    prob = O.MCProblem(5, 5, nz, 0.5, 0.5, zmax, mat, [(mua, 0.0, 0.0, n_in)],
                       n_ext=(0, 0, 0, 0, n_bot, n_top), beam=0, radius=0.05, rng=0, seed=77)
    nb, per = 8, 40000
    planes, acc = [], dict(specular=0.0, bottom=0.0, top=0.0, absorbed=0.0)
    for b in range(nb):
        tally, st = prob.run(per, p0=b * per)
        planes.append(tally.sum(axis=(1, 2)) / per)
        acc["specular"] += st.specular
        acc["bottom"] += st.escaped[4]
        acc["top"] += st.escaped[5]
        acc["absorbed"] += st.absorbed
        assert st.missed == 0 and sum(st.escaped[:4]) == 0
    N = nb * per
    for k, p in fates.items():
        sig = math.sqrt(max(p * (1 - p), 1e-12) / N)
        assert abs(acc[k] / N - p) < 4 * sig + 1e-12, (k, acc[k] / N, p)
    planes = np.array(planes)
    se = planes.std(axis=0, ddof=1) / math.sqrt(nb)
    z = np.abs(planes.mean(axis=0) - plane) / se
    assert z.max() < 4.5, z.max()
    # the reflected pass is not negligible against the statistics
    if Rb > 0:
        assert (Rb * math.exp(-mua * L) * (1 - Rt) * (1 - math.exp(-mua * dz)) / mua) > 10 * se.max()

"""Pins of the CPU oracle (CPU only).

Each oracle function is checked against something that would catch a
plausible restatement mistake:
  * golden vectors from the reference's own compiled Fortran (ran2, the photon
    walk, stokes, thermal fits) -- bit-exact;
  * the live reference build (oracle/_ref) where present -- bit-exact;
  * Philox known-answer vectors (Random123) and rocRAND's implementation;
  * closed forms: Beer-Lambert, discrete Fourier-mode decay, steady linear
    profile, constant-temperature Arrhenius, normal-incidence Fresnel,
    Brewster angle, total internal reflection;
  * distribution moments and CDFs: <cos theta> = g for both HG samplers, the
    beam radial CDFs (Kolmogorov-Smirnov);
  * invariants: packet fates sum to the packet count; the sequential rebuild
    sweep equals the single-pass removal rule on random tiny grids.
"""
import ctypes as C
import json
import math
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_ffi as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))


def hexs(a):
    return np.array([float.fromhex(x) for x in a])


# ------------------------------------------------------------------ RNG ----

def test_ran2_matches_reference_vectors():
    g = GOLD["ran2"]
    assert np.array_equal(O.ran2_seq(g["seed"], 64), hexs(g["values"]))


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_ran2_matches_live_reference_long_stream():
    s = C.c_int(-7)
    out = np.zeros(100000)
    O.ref().ref_ran2_seq(C.byref(s), 100000, out)
    assert np.array_equal(O.ran2_seq(-7, 100000), out)


KAT = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
       ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
       ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
        (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    assert O.philox(ctr, key) == expect


ROCRAND_PROBE = r'''
#include <cstdio>
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_philox4x32_10.h>
struct P : rocrand_device::philox4x32_10_engine {
    uint4 r(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
    P p; unsigned s = 12345u;
    for (int i = 0; i < 64; ++i) {
        s = s * 1664525u + 1013904223u; unsigned a = s; s = s * 1664525u + 1013904223u; unsigned b = s;
        uint4 c = make_uint4(a, b, a ^ 0x5bd1e995u, b + 7u); uint2 k = make_uint2(b * 3u, a + 11u);
        uint4 o = p.r(c, k);
        printf("%u %u %u %u %u %u %u %u %u %u\n", c.x, c.y, c.z, c.w, k.x, k.y, o.x, o.y, o.z, o.w);
    }
}
'''


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc absent")
def test_philox_matches_rocrand_implementation():
    """rocRAND's Philox4x32-10 (an independent implementation in the ROCm
    headers) on 64 pseudo-random counters/keys."""
    d = tempfile.mkdtemp()
    src, exe = os.path.join(d, "p.cpp"), os.path.join(d, "p")
    open(src, "w").write(ROCRAND_PROBE)
    r = subprocess.run(["hipcc", "-O1", "-x", "hip", "--offload-arch=gfx950", "-I/opt/rocm/include", src, "-o", exe],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("rocRAND probe did not compile: " + r.stderr[-200:])
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    n = 0
    for ln in out:
        if not ln.strip():
            continue
        v = [int(x) for x in ln.split()]
        assert O.philox(v[0:4], v[4:6]) == tuple(v[6:10])
        n += 1
    assert n == 64


def test_u01_range_and_exactness():
    L = O.lib()
    assert L.or_u01(0) == 0.5 / 2 ** 23
    assert L.or_u01(0xffffffff) == 1.0 - 0.5 / 2 ** 23
    # exactly representable in FP32
    for x in (0, 1 << 9, 0x12345678, 0xffffffff):
        u = L.or_u01(x)
        assert float(np.float32(u)) == u


# ------------------------------------------------------ photon walk ----

def test_reference_walk_matches_reference_vectors():
    g = GOLD["mc"]
    jm, esc = O.ref_mc_oracle(g["n"], g["xmax"], g["ymax"], g["zmax"], g["nphotons"], g["seed"])
    assert esc == g["escapes"]
    assert np.array_equal(jm.sum(axis=(1, 2)), hexs(g["layer_sums"]))
    assert float(jm.sum()) == float.fromhex(g["sum"])
    assert float((jm ** 2).sum()) == float.fromhex(g["sum_sq"])
    assert int((jm != 0).sum()) == g["nonzero"]


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_reference_walk_matches_live_reference_with_crater():
    """A non-homogeneous rhokap (an empty crater at the top and a dense core)
    through the reference's own tauint1: bit-exact."""
    n = 80
    rk = np.zeros((n + 2, n + 2, n + 2))
    rk[1:-1, 1:-1, 1:-1] = 680.0
    rk[n - 5:n + 1, 35:45, 35:45] = 0.0               # crater in the top layers (z index = first axis)
    rk[30:40, 30:50, 30:50] = 2000.0
    jm_o, esc_o = O.ref_mc_oracle(n, 0.03, 0.03, 0.06, 5000, -3, rhokap=rk)
    R = O.ref()
    jm = np.zeros(n ** 3)
    esc = C.c_long()
    s = C.c_int(-3)
    R.ref_mc(5000, 0.03, 0.03, 0.06, C.byref(s), 1, np.ascontiguousarray(rk), jm, C.byref(esc))
    assert np.array_equal(jm.reshape(n, n, n), jm_o)
    assert esc.value == esc_o


def _layer_expect(mu, dz, nz):
    z = np.arange(nz + 1) * dz
    return np.exp(-mu * z[:-1]) - np.exp(-mu * z[1:])


@pytest.mark.parametrize("rng", [0, 1])
def test_extended_walk_beer_lambert(rng):
    """Pure absorber, normal incidence, matched faces: per-layer tally of
    mu_a * L per packet = exp(-mu z1) - exp(-mu z2)."""
    n, npk, mu = 24, 40000, 30.0
    p = O.MCProblem(n, n, n, 0.05, 0.05, 0.05, np.zeros((n, n, n)), [[mu, 0.0, 0.9, 1.4]], beam=0,
                    radius=0.01, rng=rng, scatter=0, seed=9)
    t, st = p.run(npk)
    layer = (t * mu).sum(axis=(1, 2))[::-1] / npk
    e = _layer_expect(mu, 0.1 / n, n)
    sig = np.sqrt(e * (1 - e) / npk) + 1e-12
    assert np.max(np.abs(layer - e) / sig) < 5
    assert st.absorbed + sum(st.escaped) == npk
    assert abs(st.escaped[4] / npk - math.exp(-mu * 0.1)) < 5 * math.sqrt(math.exp(-3) / npk)


def test_extended_ran2_walk_equals_reference_walk():
    """or_mc_run with ran2, the reference beam and no Fresnel reproduces the
    reference walk (or_ref_mc) to rounding: tally * rhokap == jmean."""
    n = 40
    jo, _ = O.ref_mc_oracle(n, 0.03, 0.03, 0.06, 3000, -95648324)
    p = O.MCProblem(n, n, n, 0.03, 0.03, 0.06, np.zeros((n, n, n)), [[680.0, 0.0, 0.9, 1.38]], beam=0,
                    radius=125e-4, rng=1, scatter=1, ran2_seed=-95648324)
    t, _ = p.run(3000)
    assert np.allclose(t * 680.0, jo, rtol=1e-12, atol=1e-12)


def test_fates_sum_to_packets_with_fresnel():
    n = 16
    p = O.MCProblem(n, n, n, 0.1, 0.1, 0.1, np.zeros((n, n, n)), [[1.0, 80.0, 0.8, 1.4]],
                    n_ext=(0, 0, 0, 0, 1.33, 1.0), beam=1, radius=0.05, rng=0, seed=1)
    _, st = p.run(20000)
    assert st.absorbed + st.specular + sum(st.escaped) + st.missed == 20000
    # specular fraction is the normal-incidence Fresnel reflectance
    rsp = ((1.4 - 1.0) / 2.4) ** 2
    assert abs(st.specular / 20000 - rsp) < 5 * math.sqrt(rsp * (1 - rsp) / 20000)


# ------------------------------------------------------------- samplers ----

def test_stokes_matches_reference_vectors():
    g = GOLD["stokes"]
    L = O.lib()
    s = O.Ran2()
    L.or_ran2_init(C.byref(s), g["seed"])
    a = np.array(g["start"], dtype=float)
    nn = np.zeros(3)
    got = []
    for _ in range(200):
        L.or_stokes(a, nn, g["hgg"], C.byref(s))
        got.extend(nn)
    assert np.array_equal(np.array(got), hexs(g["dirs"]))


@pytest.mark.parametrize("gval", [0.0, 0.5, 0.9])
def test_hg_mean_cosine_is_g_rotation_sampler(gval):
    L = O.lib()
    L.or_hg_cos.restype = C.c_double
    L.or_hg_cos.argtypes = [C.c_double, C.c_double]
    xi = (np.arange(200000) + 0.5) / 200000            # stratified: the mean is an integral
    ct = np.array([L.or_hg_cos(gval, x) for x in xi])
    assert abs(ct.mean() - gval) < 2e-4
    # second moment of HG: (1 + 2 g^2) / 3
    assert abs((ct ** 2).mean() - (1 + 2 * gval ** 2) / 3) < 2e-4


def test_hg_mean_cosine_is_g_stokes():
    """Deflection cosines between successive stokes directions average to g."""
    L = O.lib()
    s = O.Ran2()
    L.or_ran2_init(C.byref(s), -99)
    a = np.array([-1.0, 0.0, 1.0, 0.0, 0.0])
    prev = np.array([0.0, 0.0, -1.0])
    nn = np.zeros(3)
    cs = []
    for _ in range(100000):
        L.or_stokes(a, nn, 0.8, C.byref(s))
        cs.append(float(prev @ nn))
        prev = nn.copy()
    cs = np.array(cs)
    assert abs(cs.mean() - 0.8) < 4 * cs.std() / math.sqrt(len(cs))


@pytest.mark.parametrize("beam,cdf", [
    (0, lambda r, R: np.clip(r * r / (R * R), 0, 1)),
    (1, lambda r, w: 1 - np.exp(-2 * r * r / (w * w)))])
def test_beam_radial_cdf(beam, cdf):
    L = O.lib()
    L.or_beam_sample.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double * 2)]
    rng = np.random.default_rng(4)
    R = 0.15
    xy = (C.c_double * 2)()
    rs = []
    for u0, u1 in rng.random((20000, 2)):
        L.or_beam_sample(beam, R, u0, u1, C.byref(xy))
        rs.append(math.hypot(xy[0], xy[1]))
    rs = np.sort(rs)
    emp = np.arange(1, len(rs) + 1) / len(rs)
    ks = np.max(np.abs(emp - cdf(rs, R)))
    assert ks < 1.63 / math.sqrt(len(rs))           # KS at 1%


def test_fresnel_closed_forms():
    L = O.lib()
    assert math.isclose(L.or_fresnel(1.4, 1.0, 1.0), (0.4 / 2.4) ** 2, rel_tol=1e-14)
    assert math.isclose(L.or_fresnel(1.0, 1.38, 1.0), (0.38 / 2.38) ** 2, rel_tol=1e-14)
    # total internal reflection beyond the critical angle (sin = 1/1.4)
    assert L.or_fresnel(1.4, 1.0, math.cos(math.asin(1 / 1.4) + 0.01)) == 1.0
    # Brewster: rp = 0, R = rs^2 / 2
    tb = math.atan(1.0 / 1.4)
    ci = math.cos(tb)
    ct = math.cos(math.asin(1.4 * math.sin(tb)))
    rs = (1.4 * ci - ct) / (1.4 * ci + ct)
    assert math.isclose(L.or_fresnel(1.4, 1.0, ci), rs * rs / 2, rel_tol=1e-12)
    assert L.or_fresnel(1.4, 1.4, 0.3) == 0.0


@pytest.mark.slow
def test_stokes_and_rotation_transport_agree():
    """The reference's stokes angles and the direction-cosine rotation give
    the same transport: per-layer tallies within 4 sigma (batch means)."""
    n, nb, per = 16, 8, 6000
    tab = [[2.0, 60.0, 0.85, 1.4]]
    res = {}
    for variant, (rng, sc) in {"rot": (0, 0), "stokes": (1, 1)}.items():
        bs = []
        for b in range(nb):
            p = O.MCProblem(n, n, n, 0.1, 0.1, 0.1, np.zeros((n, n, n)), tab, beam=0, radius=0.05, rng=rng,
                            scatter=sc, seed=100 + b, ran2_seed=-(1000 + b), stream=b)
            t, _ = p.run(per)
            bs.append(t.sum(axis=(1, 2)))
        res[variant] = np.array(bs)
    a, b = res["rot"], res["stokes"]
    se = np.sqrt(a.var(0, ddof=1) / nb + b.var(0, ddof=1) / nb)
    z = np.abs(a.mean(0) - b.mean(0)) / np.maximum(se, 1e-12)
    assert z.max() < 4


# ------------------------------------------------------------------ heat ----

def _heat(n, T, kappa, rho, cp, delt, dscale, dx, steps, jmean=None, robin=False, boil=False, Q=None):
    h = O.Heat()
    h.nx = h.ny = h.nz = n
    h.dx = h.dy = h.dz = dx
    h.delt, h.dscale = delt, dscale
    coeff = np.zeros_like(T)
    coeff[1:-1, 1:-1, 1:-1] = delt / (rho * cp)[1:-1, 1:-1, 1:-1]
    arrs = [np.array(a, dtype=float, order="C") for a in (T, kappa, rho, cp, coeff)]   # copies
    jm = np.ascontiguousarray(np.zeros((n, n, n)) if jmean is None else jmean, dtype=float)
    Qa = np.ascontiguousarray(np.zeros((n, n, n)) if Q is None else Q, dtype=float)
    h.temp, h.kappa, h.density, h.heatcap, h.coeff = [a.ctypes.data for a in arrs]
    h.jmean, h.Q = jm.ctypes.data, Qa.ctypes.data
    V = dx ** 3
    h.volumeVoxel, h.massVoxel, h.QVapor = V, rho.flat[0] * V, 2256e3 * rho.flat[0] * V
    h.boil, h.robin_top = int(boil), int(robin)
    rc = O.lib().or_heat_steps(C.byref(h), steps, None)
    return arrs[0], Qa, h.flux_out, rc


@pytest.mark.parametrize("dscale", [0.5, 1.0])
def test_heat_fourier_mode_decay(dscale):
    """Uniform medium, zero Dirichlet halo: the lowest sine mode decays by
    lam = 1 - 4 s a dt/h^2 * 3 sin^2(pi/(2(n+1))) per step."""
    n, dx = 12, 1e-4
    k, rho, cp = 0.5, 1000.0, 4000.0
    alpha = k / (rho * cp)
    delt = 0.4 * dx * dx / (6 * dscale * alpha)
    idx = np.arange(n + 2)
    s1 = np.sin(np.pi * idx / (n + 1))
    mode = s1[:, None, None] * s1[None, :, None] * s1[None, None, :]
    T = 1.0 * mode
    ones = np.ones_like(T)
    Tn, _, _, rc = _heat(n, T, k * ones, rho * ones, cp * ones, delt, dscale, dx, 50)
    lam = 1 - 4 * dscale * alpha * delt / dx ** 2 * 3 * math.sin(math.pi / (2 * (n + 1))) ** 2
    inner = (slice(1, -1),) * 3
    assert np.allclose(Tn[inner], lam ** 50 * mode[inner], rtol=1e-11, atol=1e-14)


def test_heat_linear_profile_is_steady():
    n, dx = 10, 1e-4
    z = np.arange(n + 2, dtype=float)
    T = np.broadcast_to((300 + 2.5 * z)[:, None, None], (n + 2, n + 2, n + 2)).copy()
    ones = np.ones_like(T)
    Tn, _, _, _ = _heat(n, T, 0.5 * ones, 1000 * ones, 4000 * ones, 0.01, 0.5, dx, 200)
    assert np.allclose(Tn, T, rtol=0, atol=1e-9)


def test_heat_energy_balance():
    """Uniform rho c: heat content change = source energy + boundary inflow."""
    n, dx = 14, 1e-4
    ones = np.ones((n + 2,) * 3)
    T = 300.0 * ones
    jm = np.zeros((n, n, n))
    jm[5:9, 5:9, 5:9] = 5e8
    rho, cp = 1000.0, 4000.0
    delt = 0.01                          # 0.75 of the explicit limit dx^2 / (6 alpha)
    Tn, _, flux, _ = _heat(n, T, 0.5 * ones, rho * ones, cp * ones, delt, 1.0, dx, 300, jmean=jm)
    V = dx ** 3
    dE = ((Tn - T)[1:-1, 1:-1, 1:-1] * rho * cp * V).sum()
    src = jm.sum() * V * delt * 300
    assert abs(dE - (src - flux)) < 1e-9 * src
    assert flux > 0


def test_heat_boiling_clamp_closed_form():
    """A voxel at 373 K with latent heat left stays at 373 K and banks the
    step's energy in Q (3dFD.f90:169-172)."""
    n, dx = 3, 1e-4
    ones = np.ones((n + 2,) * 3)
    T = 373.0 * ones
    jm = np.zeros((n, n, n))
    jm[1, 1, 1] = 1e9
    rho, cp, delt = 1000.0, 4000.0, 1e-3
    Tn, Q, _, _ = _heat(n, T, 0.5 * ones, rho * ones, cp * ones, delt, 0.5, dx, 1, jmean=jm, boil=True)
    assert Tn[2, 2, 2] == 373.0
    assert math.isclose(Q[1, 1, 1], 1e9 * delt * dx ** 3, rel_tol=1e-14)


def test_heat_negative_temperature_flagged():
    n = 3
    ones = np.ones((n + 2,) * 3)
    T = 300.0 * ones
    T[2, 2, 2] = -1.0
    _, _, _, rc = _heat(n, T, 0.5 * ones, 1000 * ones, 4000 * ones, 1e-3, 0.5, 1e-4, 1)
    assert rc == -1


# ---------------------------------------------------- damage, rebuild ----

def test_arrhenius_constant_temperature_closed_form_and_chain():
    n = 4
    T = np.full((n + 2,) * 3, 340.0)
    T[1, 1, 1] = 350.0                     # faster voxel crosses two thresholds in one call
    rk = np.ones_like(T)
    a = O.Arrh()
    a.A, a.dE, a.R, a.Tlo, a.Thi = 3.1e98, 6.3e5, 8.314, 316.0, 373.0
    a.thr = (C.c_double * 3)(0.53, 1.0, 1e4)
    om = np.zeros((n, n, n))
    thr = np.zeros((3, n, n, n))
    dt = 0.7
    rate = 3.1e98 * math.exp(-6.3e5 / (8.314 * 340.0))
    calls = 40
    for c in range(calls):
        O.lib().or_arrhenius(C.byref(a), n, n, n, T, rk, dt, 10.0 * (c + 1), om, thr)
    assert math.isclose(om[1, 1, 1], calls * dt * rate, rel_tol=1e-12)
    first = math.ceil(0.53 / (dt * rate))
    assert thr[0, 1, 1, 1] == 10.0 * first
    # the fast voxel: if one call passes 0.53 and 1.0 together, 1.0 is
    # recorded on the next call (elseif chain, 3dFD.f90:455-461)
    r2 = 3.1e98 * math.exp(-6.3e5 / (8.314 * 350.0))
    c1 = math.ceil(0.53 / (dt * r2))
    c2 = math.ceil(1.0 / (dt * r2))
    assert thr[0, 0, 0, 0] == 10.0 * c1
    assert thr[1, 0, 0, 0] == 10.0 * (c2 if c2 > c1 else c1 + 1)


def test_arrhenius_window_edges():
    n = 1
    a = O.Arrh()
    a.A, a.dE, a.R, a.Tlo, a.Thi = 3.1e98, 6.3e5, 8.314, 316.0, 373.0
    a.thr = (C.c_double * 3)(0.53, 1.0, 1e4)
    for Tval, inside in ((316.0, True), (372.999, True), (373.0, False), (315.999, False)):
        T = np.full((3, 3, 3), Tval)
        om = np.zeros((1, 1, 1))
        O.lib().or_arrhenius(C.byref(a), n, n, n, T, np.ones_like(T), 1.0, 1.0, om, np.zeros((3, 1, 1, 1)))
        assert (om[0, 0, 0] > 0) == inside


def _seq_rebuild_zero_pattern(T, rk, ablate):
    n = T.shape[0] - 2
    r = O.Rebuild()
    r.nx = r.ny = r.nz = n
    r.mode, r.ablate_K, r.delt = 1, ablate, 1e-3
    arrs = {k: np.ascontiguousarray(v) for k, v in dict(
        rhokap=rk.copy(), density=np.full_like(T, 1000.0), heatcap=np.full_like(T, 4000.0),
        kappa=np.full_like(T, 0.5), coeff=np.zeros_like(T), alpha=np.zeros_like(T)).items()}
    W = np.full((n, n, n), 0.75)
    Q = np.zeros((n, n, n))
    for k in arrs:
        setattr(r, k, arrs[k].ctypes.data)
    Ta = np.ascontiguousarray(T)
    r.W, r.Q, r.temp = W.ctypes.data, Q.ctypes.data, Ta.ctypes.data
    O.lib().or_rebuild(C.byref(r))
    return arrs["rhokap"][1:-1, 1:-1, 1:-1] == 0


@pytest.mark.parametrize("n", [3, 4, 5, 7])
def test_removal_rule_equals_sequential_sweep(n):
    """Brute force over many random tiny grids: the single-pass rule
    Z = B | (U & B(-x) & B(-y) & B(-z)) reproduces the in-place sweep of
    setupThermalCoeff exactly."""
    rng = np.random.default_rng(n)
    for trial in range(300):
        rk = np.zeros((n + 2,) * 3)
        rk[1:-1, 1:-1, 1:-1] = np.where(rng.random((n, n, n)) < rng.uniform(0.2, 0.8), 0.0, 680.0)
        T = 300 + 200 * rng.random((n + 2,) * 3)
        mask = np.zeros((n, n, n), np.uint8)
        O.lib().or_removal_mask(n, n, n, np.ascontiguousarray(T), np.ascontiguousarray(rk), 400.0, mask)
        assert np.array_equal(mask.astype(bool), _seq_rebuild_zero_pattern(T, rk, 400.0)), trial


def test_removal_hand_case():
    """A cool tissue voxel whose +x,+y,+z neighbours are empty and whose
    -x,-y,-z neighbours are ablated in the same sweep is removed; with one
    tissue +neighbour it stays."""
    n = 3
    rk = np.zeros((n + 2,) * 3)
    rk[1:-1, 1:-1, 1:-1] = 680.0
    T = np.full((n + 2,) * 3, 300.0)
    c = (2, 2, 2)
    for d in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
        rk[c[0] + d[0], c[1] + d[1], c[2] + d[2]] = 0.0          # +neighbours empty
        T[c[0] - d[0], c[1] - d[1], c[2] - d[2]] = 500.0         # -neighbours ablate
    z = _seq_rebuild_zero_pattern(T, rk, 400.0)
    assert z[1, 1, 1]
    rk[3, 2, 2] = 680.0
    z = _seq_rebuild_zero_pattern(T, rk, 400.0)
    assert not z[1, 1, 1]


def test_thermal_fits_match_reference_vectors():
    g = GOLD["fits"]
    L = O.lib()
    out = hexs(g["out"]).reshape(15, 6)
    for i, (T, W, Q, rho) in enumerate(zip(g["T"], g["W"], g["Q"], g["rho"])):
        got = [L.or_airThermalCond(T), L.or_airDensity(T), L.or_getWaterContent(Q, W, g["qvapor"]),
               L.or_getSkinDensity(W), L.or_getSkinHeatCap(W), L.or_getSkinThermalCond(W, rho)]
        assert got == list(out[i]), i


def test_water_content_closed_form():
    L = O.lib()
    assert L.or_getWaterContent(0.0, 0.75, 1.0) == 0.75
    assert L.or_getWaterContent(0.5, 0.75, 1.0) == 0.375
    assert L.or_getWaterContent(2.0, 0.75, 1.0) == 0.0              # clamped at zero
    assert L.or_getWaterContent(0.1, 0.2, 1.0) == 0.2               # never rises above the current W


def test_init_thermal_defaults():
    """initThermalCoeff at input.params: delt = 1/(alpha sum 1/dx^2), the
    values quoted for the reference default (delt 1.628e-4 s)."""
    t = O.ThermalInit()
    O.lib().or_init_thermal(80, 0.03, 0.03, 0.06, 80, 80, 80, 400.0, 70.0, C.byref(t))
    assert math.isclose(t.dx, 0.06e-2 / 82, rel_tol=1e-15)
    assert math.isclose(t.delt, 1.628e-4, rel_tol=2e-3)
    assert math.isclose(t.density, 1000 / (0.75 + 0.649 * 0.25), rel_tol=1e-15)
    assert math.isclose(t.coeff, t.delt / (t.density * t.heatcap), rel_tol=1e-12)
    assert math.isclose(t.pulseLength, 0.4 * 81 / 70, rel_tol=1e-15)


def test_pulse_gaussian_peak_and_state_machine():
    L = O.lib()
    p = O.Pulse()
    p.power, p.pulseLength, p.type = 70.0, 0.4628, 1
    fact = 2 * math.sqrt(2 * math.log(2))
    p.time = fact * p.pulseLength
    assert math.isclose(L.or_getPwr(C.byref(p)), 70.0, rel_tol=1e-15)
    # tophat: on until pulseCount reaches realPulseLength, then off (one
    # pulse).  With an exact binary step the laser is on for
    # realPulseLength/delt + 1 steps: the off test runs before the count
    # advances (3dFD.f90:199-213).
    p = O.Pulse()
    p.power, p.realPulseLength, p.type = 5.0, 1.0, 0
    p.laser_flag, p.laserOn, p.pulsesToDo, p.repetitionRate_1 = 1, 1.0, 1, 1e9
    n_on = 0
    for _ in range(30):
        n_on += p.laserOn > 0
        L.or_pulse_advance(C.byref(p), 0.125)
    assert p.laser_flag == 0 and p.pulsesDone == 1 and n_on == 9

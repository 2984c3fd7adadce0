"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and, where it was
built, the reference Fortran compiled from its own sources (oracle/_ref/libref.so).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class Ran2(C.Structure):
    _fields_ = [("idum", C.c_int32), ("idum2", C.c_int32), ("iy", C.c_int32),
                ("iv", C.c_int32 * 32)]


class Grid(C.Structure):
    _fields_ = [("nxg", C.c_int), ("nyg", C.c_int), ("nzg", C.c_int),
                ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double)]


class Material(C.Structure):
    _fields_ = [("mua", C.c_double), ("mus", C.c_double), ("g", C.c_double), ("n", C.c_double)]


class MC(C.Structure):
    _fields_ = [("grid", Grid), ("mut", C.c_void_p), ("mat", C.c_void_p), ("table", C.c_void_p),
                ("nmat", C.c_int), ("n_ext", C.c_double * 6), ("beam", C.c_int),
                ("beam_radius", C.c_double), ("rng", C.c_int), ("scatter", C.c_int),
                ("seed", C.c_uint64), ("stream", C.c_uint32), ("ran2_seed", C.c_int32)]


class MCStats(C.Structure):
    _fields_ = [("absorbed", C.c_double), ("specular", C.c_double), ("escaped", C.c_double * 6),
                ("missed", C.c_double), ("crossings", C.c_double), ("interactions", C.c_double)]


class Heat(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("delt", C.c_double), ("dscale", C.c_double),
                ("temp", C.c_void_p), ("kappa", C.c_void_p), ("density", C.c_void_p),
                ("heatcap", C.c_void_p), ("coeff", C.c_void_p), ("Q", C.c_void_p),
                ("jmean", C.c_void_p), ("volumeVoxel", C.c_double), ("massVoxel", C.c_double),
                ("QVapor", C.c_double), ("boil", C.c_int), ("robin_top", C.c_int),
                ("h_top", C.c_double), ("T_amb", C.c_double), ("flux_out", C.c_double),
                ("rhokap", C.c_void_p), ("removed_inert", C.c_int)]


class Arrh(C.Structure):
    _fields_ = [("A", C.c_double), ("dE", C.c_double), ("R", C.c_double),
                ("Tlo", C.c_double), ("Thi", C.c_double), ("thr", C.c_double * 3)]


class Rebuild(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("mode", C.c_int),
                ("ablate_K", C.c_double), ("delt", C.c_double), ("mu_water", C.c_double),
                ("mu_protein", C.c_double), ("QVapor", C.c_double),
                ("rhokap", C.c_void_p), ("density", C.c_void_p), ("heatcap", C.c_void_p),
                ("kappa", C.c_void_p), ("coeff", C.c_void_p), ("alpha", C.c_void_p),
                ("W", C.c_void_p), ("Q", C.c_void_p), ("temp", C.c_void_p)]


class Pulse(C.Structure):
    _fields_ = [("power", C.c_double), ("pulseLength", C.c_double), ("realPulseLength", C.c_double),
                ("time", C.c_double), ("pulseCount", C.c_double), ("repetitionCount", C.c_double),
                ("repetitionRate_1", C.c_double), ("laserOn", C.c_double),
                ("laser_flag", C.c_int), ("pulseFlag", C.c_int), ("pulsesDone", C.c_int),
                ("pulsesToDo", C.c_int), ("type", C.c_int)]


class ThermalInit(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "dx", "dy", "dz", "delt", "alpha", "kappa_skin", "density", "heatcap", "coeff",
        "kappa_air", "alpha_air_top", "volumeVoxel", "massVoxel", "QVapor", "pulseLength")]


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        L = C.CDLL(ORACLE_SO)
        L.or_ran2.restype = C.c_double
        L.or_u01.restype = C.c_double
        L.or_u01.argtypes = [C.c_uint32]
        L.or_fresnel.restype = C.c_double
        L.or_fresnel.argtypes = [C.c_double, C.c_double, C.c_double]
        L.or_ref_mc.restype = C.c_long
        L.or_ref_mc.argtypes = [C.POINTER(Grid), _dp, C.c_long, C.POINTER(Ran2), _dp]
        L.or_mc_run.restype = C.c_int
        L.or_mc_run.argtypes = [C.POINTER(MC), C.c_uint64, C.c_uint64, _dp, C.POINTER(MCStats), C.c_int]
        L.or_heat_steps.restype = C.c_int
        L.or_heat_steps.argtypes = [C.POINTER(Heat), C.c_int, C.c_void_p]
        L.or_arrhenius.argtypes = [C.POINTER(Arrh), C.c_int, C.c_int, C.c_int, _dp, _dp,
                                   C.c_double, C.c_double, _dp, _dp]
        L.or_removal_mask.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_double, _u8p]
        for f in ("or_airThermalCond", "or_airDensity", "or_getSkinDensity", "or_getSkinHeatCap"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_double]
        L.or_getWaterContent.restype = C.c_double
        L.or_getWaterContent.argtypes = [C.c_double, C.c_double, C.c_double]
        L.or_getSkinThermalCond.restype = C.c_double
        L.or_getSkinThermalCond.argtypes = [C.c_double, C.c_double]
        L.or_getPwr.restype = C.c_double
        L.or_getPwr.argtypes = [C.POINTER(Pulse)]
        L.or_pulse_advance.argtypes = [C.POINTER(Pulse), C.c_double]
        L.or_init_thermal.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                      C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(ThermalInit)]
        L.or_stokes.argtypes = [_dp, _dp, C.c_double, C.POINTER(Ran2)]
        L.or_mk_cdf.argtypes = [_dp, _dp, C.c_int]
        L.or_search_1d.argtypes = [C.c_int, _dp, C.c_double]
        L.or_search_2d.argtypes = [C.c_int, _dp, C.c_double]
        L.or_lin_inter_1d.argtypes = [_dp, _dp, C.c_double, C.c_int, C.c_int]
        L.or_lin_inter_1d.restype = C.c_double
        L.or_lin_inter_2d.argtypes = [_dp, C.c_double, C.c_int, C.c_int]
        L.or_lin_inter_2d.restype = C.c_double
        L.or_sample.argtypes = [_dp, C.c_int, _dp, C.POINTER(Ran2)]
        L.or_sample.restype = C.c_double
        L.or_sample_u.argtypes = [_dp, C.c_int, _dp, C.c_double]
        L.or_sample_u.restype = C.c_double
        _lib = L
    return _lib


_ref = None


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    """The reference's own Fortran (compiled by `make -C oracle ref`)."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_SO)
        R.ref_ran2_seq.argtypes = [C.POINTER(C.c_int), C.c_int, _dp]
        R.ref_grid_n.argtypes = [C.POINTER(C.c_int)]
        R.ref_mc.argtypes = [C.c_long, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int),
                             C.c_int, _dp, _dp, C.POINTER(C.c_long)]
        R.ref_stokes.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_int), _dp, _dp]
        R.ref_fits.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, _dp, _dp]
        _ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        R.ref_mk_cdf.argtypes = [C.c_int, _dp, _dp]
        R.ref_sample.argtypes = [C.c_int, C.c_int, _dp, _dp, C.POINTER(C.c_int), _dp]
        R.ref_search.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _ip]
        R.ref_lin_inter.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _ip, _ip, _dp]
        _ref = R
    return _ref


# ----------------------------------------------------------------- helpers --

def philox(ctr, key):
    L = lib()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.or_philox4x32_10(c, k, o)
    return tuple(o)


def ran2_seq(iseed, n):
    L = lib()
    s = Ran2()
    L.or_ran2_init(C.byref(s), C.c_int32(iseed))
    return np.array([L.or_ran2(C.byref(s)) for _ in range(n)])


def ref_mc_oracle(n, xmax, ymax, zmax, nphotons, iseed, rhokap=None, kappa=680.0):
    """or_ref_mc on an n^3 grid; returns (jmean[z,y,x] C-order view of col-major, escapes)."""
    L = lib()
    g = Grid(n, n, n, xmax, ymax, zmax)
    if rhokap is None:
        rhokap = np.zeros((n + 2, n + 2, n + 2))
        rhokap[1:-1, 1:-1, 1:-1] = kappa
    s = Ran2()
    L.or_ran2_init(C.byref(s), C.c_int32(iseed))
    jm = np.zeros(n * n * n)
    esc = L.or_ref_mc(C.byref(g), np.ascontiguousarray(rhokap, dtype=np.float64), nphotons,
                      C.byref(s), jm)
    return jm.reshape(n, n, n), esc


class MCProblem:
    """Extended-physics MC problem for or_mc_run (arrays indexed [z, y, x])."""

    def __init__(self, nx, ny, nz, xmax, ymax, zmax, mat, table, mut=None,
                 n_ext=(0, 0, 0, 0, 0, 0), beam=0, radius=0.1, rng=0, scatter=0,
                 seed=0, stream=0, ran2_seed=-95648324):
        self.shape = (nz, ny, nx)
        self.mat = np.ascontiguousarray(mat, dtype=np.uint8).reshape(nz, ny, nx)
        self.table_np = np.asarray(table, dtype=np.float64).reshape(-1, 4)
        if mut is None:
            mut = self.table_np[self.mat, 0] + self.table_np[self.mat, 1]
        self.mut = np.ascontiguousarray(mut, dtype=np.float64)
        self.table = (Material * len(self.table_np))(*[Material(*r) for r in self.table_np])
        self.m = MC()
        self.m.grid = Grid(nx, ny, nz, xmax, ymax, zmax)
        self.m.mut = self.mut.ctypes.data
        self.m.mat = self.mat.ctypes.data
        self.m.table = C.cast(self.table, C.c_void_p)
        self.m.nmat = len(self.table_np)
        self.m.n_ext = (C.c_double * 6)(*n_ext)
        self.m.beam = beam
        self.m.beam_radius = radius
        self.m.rng = rng
        self.m.scatter = scatter
        self.m.seed = seed
        self.m.stream = stream
        self.m.ran2_seed = ran2_seed

    def run(self, npk, p0=0, nthreads=0):
        tally = np.zeros(int(np.prod(self.shape)))
        st = MCStats()
        rc = lib().or_mc_run(C.byref(self.m), p0, npk, tally, C.byref(st), nthreads)
        assert rc == 0
        return tally.reshape(self.shape), st


# ------------------------------------------------------ spectral helpers --

def spectrum_table(x, y):
    """(len, 2) column-major table as the reference stores array(len, 2)."""
    return np.ascontiguousarray(np.concatenate([np.asarray(x, float), np.asarray(y, float)]))


def mk_cdf(table, cdf_last=0.0):
    n = table.size // 2
    cdf = np.zeros(n)
    cdf[-1] = cdf_last
    lib().or_mk_cdf(table, cdf, n)
    return cdf


def sample_seq(table, cdf, seed, ns):
    L = lib()
    s = Ran2()
    L.or_ran2_init(C.byref(s), C.c_int32(seed))
    return np.array([L.or_sample(table, table.size // 2, cdf, C.byref(s)) for _ in range(ns)])

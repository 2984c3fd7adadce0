"""ctypes binding of libtam.so (the C-ABI in include/tam.h).

The HIP library is the only implementation: if libtam.so is missing this
module raises, it never falls back to a CPU path.

Arrays are numpy, C-order with shape (nz, ny, nx) for interior fields and
(nz+2, ny+2, nx+2) for halo fields -- the same memory layout as the
reference's column-major (x, y, z) Fortran arrays (src/iarray.f90:8-12).
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libtam.so")

TAM_OK = 0
BEAM_TOPHAT, BEAM_GAUSSIAN = 0, 1
THERMAL_WATER, THERMAL_TABLE = 0, 1
FACE_XLO, FACE_XHI, FACE_YLO, FACE_YHI, FACE_ZLO, FACE_ZHI = range(6)
(FIELD_TEMP, FIELD_JMEAN, FIELD_TALLY, FIELD_RHOKAP, FIELD_TISSUE, FIELD_THRESTIME, FIELD_WATER,
 FIELD_Q, FIELD_KAPPA, FIELD_DENSITY, FIELD_HEATCAP, FIELD_COEFF, FIELD_MAT) = range(13)
HALO_FIELDS = {FIELD_TEMP, FIELD_RHOKAP, FIELD_KAPPA, FIELD_DENSITY, FIELD_HEATCAP, FIELD_COEFF}

EXPORTED = ("tam_abi_version", "tam_create", "tam_destroy", "tam_last_error", "tam_set_optics",
            "tam_set_thermal", "tam_mc_run", "tam_normalise", "tam_heat_run", "tam_arrhenius",
            "tam_rebuild", "tam_download", "tam_upload", "tam_energy", "tam_get_timing",
            "tam_synchronize", "tam_set_heat_launch", "tam_tune_heat", "tam_set_mc_launch", "tam_tune_mc",
            "tam_set_mc_box", "tam_get_mc_box",
            "tam_copy_bandwidth", "tam_atomic_rate", "tam_struct_size", "tam_pulse_advance", "tam_cycle",
            "tam_tally_blocks", "tam_batch_create", "tam_batch_destroy", "tam_batch_last_error",
            "tam_batch_mc_run", "tam_batch_heat_run", "tam_batch_set_launch", "tam_batch_tune", "tam_batch_last_ms",
            "tam_spectrum_cdf", "tam_spectrum_sample", "tam_spectrum_interp", "tam_tally_blocks_noalloc")


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
        ("dx_m", C.c_double), ("dy_m", C.c_double), ("dz_m", C.c_double),
        ("delt", C.c_double), ("dscale", C.c_double),
        ("volume_voxel", C.c_double), ("mass_voxel", C.c_double), ("qvapor", C.c_double),
        ("t_ref", C.c_double), ("boil", C.c_int32), ("robin_top", C.c_int32),
        ("h_top", C.c_double), ("t_amb", C.c_double),
        ("thermal_mode", C.c_int32), ("mu_water", C.c_double), ("mu_protein", C.c_double),
        ("ablate_K", C.c_double),
        ("arr_A", C.c_double), ("arr_dE", C.c_double), ("arr_R", C.c_double),
        ("arr_Tlo", C.c_double), ("arr_Thi", C.c_double), ("arr_thr", C.c_double * 3),
        ("damage_every_step", C.c_int32),
        ("beam", C.c_int32), ("beam_radius", C.c_double),
        ("n_ext_zlo", C.c_double), ("n_ext_zhi", C.c_double),
        ("device", C.c_int32), ("heat_fp64", C.c_int32), ("removed_inert", C.c_int32),
    ]


class Material(C.Structure):
    _fields_ = [("mua", C.c_double), ("mus", C.c_double), ("g", C.c_double), ("n", C.c_double)]


class MCStats(C.Structure):
    _fields_ = [("packets", C.c_double), ("absorbed", C.c_double), ("specular", C.c_double),
                ("escaped", C.c_double * 6), ("missed", C.c_double), ("crossings", C.c_double),
                ("interactions", C.c_double), ("kernel_ms", C.c_double), ("killed", C.c_double),
                ("tally_adds", C.c_double), ("box_adds", C.c_double)]

    def fractions(self):
        n = self.packets
        return {"absorbed": self.absorbed / n, "specular": self.specular / n,
                "diffuse_reflected": self.escaped[FACE_ZHI] / n,
                "transmitted": self.escaped[FACE_ZLO] / n,
                "lateral": sum(self.escaped[f] for f in range(4)) / n,
                "missed": self.missed / n}


class Pulse(C.Structure):
    _fields_ = [("time", C.c_double), ("pulseCount", C.c_double), ("repetitionCount", C.c_double),
                ("repetitionRate_1", C.c_double), ("realPulseLength", C.c_double),
                ("laserOn", C.c_double), ("laser_flag", C.c_int32), ("pulsesDone", C.c_int32),
                ("pulsesToDo", C.c_int32), ("use_pulse", C.c_int32)]


class Timing(C.Structure):
    _fields_ = [("mc_ms", C.c_double), ("normalise_ms", C.c_double), ("heat_ms", C.c_double),
                ("arrhenius_ms", C.c_double), ("rebuild_ms", C.c_double),
                ("heat_steps", C.c_double), ("heat_step_ms", C.c_double),
                ("class_mode", C.c_double), ("tile", C.c_double), ("kz", C.c_double),
                ("cycle_ms", C.c_double)]


class TamError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tam error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libtam.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built: run __graft_entry__.build() "
                          "(or python tissue-ablation-mc_amd/build.py)")
    L = C.CDLL(LIB_PATH)
    vp, dp = C.c_void_p, C.c_void_p
    L.tam_abi_version.restype = C.c_int
    L.tam_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.tam_destroy.argtypes = [vp]
    L.tam_destroy.restype = None
    L.tam_last_error.argtypes = [vp]
    L.tam_last_error.restype = C.c_char_p
    L.tam_set_optics.argtypes = [vp, vp, C.POINTER(Material), C.c_int32, dp]
    L.tam_set_thermal.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.tam_mc_run.argtypes = [vp, C.c_int64, C.c_uint64, C.c_uint32, C.c_int32, C.POINTER(MCStats)]
    L.tam_normalise.argtypes = [vp, C.c_double]
    L.tam_heat_run.argtypes = [vp, C.c_int32, C.POINTER(Pulse)]
    L.tam_arrhenius.argtypes = [vp, C.c_double, C.c_double]
    L.tam_rebuild.argtypes = [vp]
    L.tam_download.argtypes = [vp, C.c_int32, dp]
    L.tam_upload.argtypes = [vp, C.c_int32, dp]
    L.tam_energy.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.tam_get_timing.argtypes = [vp, C.POINTER(Timing)]
    L.tam_synchronize.argtypes = [vp]
    L.tam_set_heat_launch.argtypes = [vp, C.c_int32, C.c_int32]
    L.tam_set_mc_launch.argtypes = [vp, C.c_int32, C.c_int32]
    L.tam_set_mc_box.argtypes = [vp, C.c_int32]
    L.tam_get_mc_box.argtypes = [vp, C.POINTER(C.c_int32)]
    L.tam_copy_bandwidth.argtypes = [vp, C.c_int64, C.c_int32, C.POINTER(C.c_double)]
    L.tam_atomic_rate.argtypes = [vp, C.c_int64, C.c_int64, C.POINTER(C.c_double)]
    L.tam_tally_blocks.argtypes = [vp, C.c_int32, dp]
    L.tam_tally_blocks_noalloc.argtypes = [vp, C.c_int32, dp]
    L.tam_batch_create.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(vp)]
    L.tam_batch_destroy.argtypes = [vp]
    L.tam_batch_destroy.restype = None
    L.tam_batch_last_error.argtypes = [vp]
    L.tam_batch_last_error.restype = C.c_char_p
    L.tam_batch_mc_run.argtypes = [vp, C.POINTER(C.c_int64), C.c_uint64, C.c_uint32, C.c_int32, C.POINTER(MCStats)]
    L.tam_batch_heat_run.argtypes = [vp, C.c_int32]
    L.tam_batch_set_launch.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    L.tam_batch_tune.argtypes = [vp, C.POINTER(C.c_int64), C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.tam_batch_last_ms.argtypes = [vp]
    L.tam_batch_last_ms.restype = C.c_double
    L.tam_struct_size.argtypes = [C.c_int32]
    L.tam_struct_size.restype = C.c_int64
    L.tam_pulse_advance.argtypes = [C.POINTER(Pulse), C.c_double]
    L.tam_pulse_advance.restype = None
    L.tam_tune_mc.argtypes = [vp, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_double)]
    L.tam_cycle.argtypes = [vp, C.c_int64, C.c_uint64, C.c_uint32, C.c_double, C.c_int32, C.c_double,
                            C.POINTER(Pulse), C.POINTER(MCStats)]
    L.tam_tune_heat.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                C.POINTER(C.c_double)]
    L.tam_spectrum_cdf.argtypes = [dp, C.c_int32, dp]
    L.tam_spectrum_sample.argtypes = [C.c_int32, dp, dp, C.c_int32, C.c_int64, C.c_uint64, C.c_uint32, dp]
    L.tam_spectrum_interp.argtypes = [C.c_int32, dp, C.c_int32, dp, C.c_int64, dp]
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != shape:
        a = a.reshape(shape)
    return a


class TissueGrid:
    """One tissue grid resident on one GPU (a tam_ctx)."""

    def __init__(self, params: Params):
        self.L = load()
        self.p = params
        self.shape = (params.nz, params.ny, params.nx)
        self.hshape = (params.nz + 2, params.ny + 2, params.nx + 2)
        h = C.c_void_p()
        rc = self.L.tam_create(C.byref(params), C.byref(h))
        if rc != TAM_OK:
            raise TamError(rc, self.L.tam_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.tam_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != TAM_OK:
            raise TamError(rc, self.L.tam_last_error(self.h).decode())

    def set_optics(self, mat, table, rhokap=None):
        mat = np.ascontiguousarray(mat, dtype=np.uint8).reshape(self.shape)
        tab = (Material * len(table))(*[Material(*map(float, t)) for t in table])
        rk = None if rhokap is None else _f64(rhokap, self.hshape)
        self._keep_optics = (mat, rk)
        self._chk(self.L.tam_set_optics(self.h, _ptr(mat), tab, len(table), _ptr(rk)))

    def set_thermal(self, temp, kappa, density, heatcap, Q=None, W=None):
        t, k, d, c = (_f64(x, self.hshape) for x in (temp, kappa, density, heatcap))
        q = None if Q is None else _f64(Q, self.shape)
        w = None if W is None else _f64(W, self.shape)
        self._chk(self.L.tam_set_thermal(self.h, _ptr(t), _ptr(k), _ptr(d), _ptr(c), _ptr(q), _ptr(w)))

    def mc_run(self, npackets, seed, stream=0, accumulate=False):
        st = MCStats()
        self._chk(self.L.tam_mc_run(self.h, int(npackets), int(seed), int(stream), int(bool(accumulate)),
                                    C.byref(st)))
        return st

    def cycle(self, npackets, seed, stream, scale, loops, arr_delt=0.0, pulse=None, stats=True):
        """One coupling cycle as a graph replay (transport, normalise, `loops`
        heat steps, optional Arrhenius, rebuild).  Returns the transport
        counters when stats is true (the call then waits for the GPU)."""
        st = MCStats() if stats else None
        self._chk(self.L.tam_cycle(self.h, int(npackets), int(seed), int(stream), float(scale), int(loops),
                                   float(arr_delt), None if pulse is None else C.byref(pulse),
                                   None if st is None else C.byref(st)))
        return st

    def normalise(self, scale):
        self._chk(self.L.tam_normalise(self.h, float(scale)))

    def heat_run(self, loops, pulse=None):
        self._chk(self.L.tam_heat_run(self.h, int(loops), None if pulse is None else C.byref(pulse)))

    def arrhenius(self, delt, time):
        self._chk(self.L.tam_arrhenius(self.h, float(delt), float(time)))

    def rebuild(self):
        self._chk(self.L.tam_rebuild(self.h))

    def download(self, field):
        if field in HALO_FIELDS:
            shape = self.hshape
        elif field == FIELD_THRESTIME:
            shape = (3,) + self.shape
        else:
            shape = self.shape
        out = np.empty(shape, dtype=np.float64)
        self._chk(self.L.tam_download(self.h, field, _ptr(out)))
        return out

    def upload(self, field, arr):
        if field in HALO_FIELDS:
            shape = self.hshape
        elif field == FIELD_THRESTIME:
            shape = (3,) + self.shape
        else:
            shape = self.shape
        a = _f64(arr, shape)
        self._chk(self.L.tam_upload(self.h, field, _ptr(a)))

    def energy(self):
        hc, bo = C.c_double(), C.c_double()
        self._chk(self.L.tam_energy(self.h, C.byref(hc), C.byref(bo)))
        return hc.value, bo.value

    def timing(self):
        t = Timing()
        self._chk(self.L.tam_get_timing(self.h, C.byref(t)))
        return t

    def tally_blocks(self, k):
        """k^3 block sums of the transport tally, shape (nz/k, ny/k, nx/k)."""
        nz, ny, nx = self.shape
        out = np.empty((nz // k, ny // k, nx // k), dtype=np.float64)
        self._chk(self.L.tam_tally_blocks(self.h, int(k), _ptr(out)))
        return out

    def synchronize(self):
        self._chk(self.L.tam_synchronize(self.h))

    def set_heat_launch(self, tile, kz):
        self._chk(self.L.tam_set_heat_launch(self.h, int(tile), int(kz)))

    def copy_bandwidth(self, nbytes=1 << 30, reps=10):
        """Measured device copy bandwidth, GB/s (bytes read + written)."""
        g = C.c_double()
        self._chk(self.L.tam_copy_bandwidth(self.h, int(nbytes), int(reps), C.byref(g)))
        return g.value

    def atomic_rate(self, count=1 << 28, slots=8 << 20):
        """Measured FP64 atomic adds per second (random slots, no return)."""
        r = C.c_double()
        self._chk(self.L.tam_atomic_rate(self.h, int(count), int(slots), C.byref(r)))
        return r.value

    def set_mc_launch(self, blocks_per_cu, rng_batch=32):
        self._chk(self.L.tam_set_mc_launch(self.h, int(blocks_per_cu), int(rng_batch)))

    def set_mc_box(self, enable, fp64=False):
        """Block-private LDS tally box on the beam axis on (default) or off;
        fp64 keeps FP64 sums in it (a quarter of the voxels) instead of 16-bit
        fixed point."""
        self._chk(self.L.tam_set_mc_box(self.h, (2 if fp64 else 1) if enable else 0))

    def mc_box(self):
        """The LDS box of the last transport launch that placed one:
        ((x0, y0, z0), (nx, ny, nz)), 0-based interior voxel indices."""
        g = (C.c_int32 * 6)()
        self._chk(self.L.tam_get_mc_box(self.h, g))
        return tuple(g[:3]), tuple(g[3:])

    def tune_mc(self, sample=2_000_000):
        """Autotune the transport launch; returns (blocks per CU, random-phase
        batch, packets/s)."""
        b, t, r = C.c_int32(), C.c_int32(), C.c_double()
        self._chk(self.L.tam_tune_mc(self.h, int(sample), C.byref(b), C.byref(t), C.byref(r)))
        return b.value, t.value, r.value

    def tune_heat(self, reps=5):
        """Autotune the stencil launch shape; returns (tile, kz, ms per step)."""
        t, k, ms = C.c_int32(), C.c_int32(), C.c_double()
        self._chk(self.L.tam_tune_heat(self.h, int(reps), C.byref(t), C.byref(k), C.byref(ms)))
        return t.value, k.value, ms.value


class Batch:
    """Independent TissueGrids run together: one transport launch for all,
    one stencil launch per step for all (tam_batch_*)."""

    def __init__(self, grids):
        self.L = load()
        self.grids = list(grids)
        arr = (C.c_void_p * len(self.grids))(*[g.h for g in self.grids])
        h = C.c_void_p()
        rc = self.L.tam_batch_create(arr, len(self.grids), C.byref(h))
        if rc != TAM_OK:
            raise TamError(rc, self.L.tam_batch_last_error(None).decode())
        self.h = h

    def _chk(self, rc):
        if rc != TAM_OK:
            raise TamError(rc, self.L.tam_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.tam_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mc_run(self, npackets, seed, stream0=0, accumulate=False):
        n = len(self.grids)
        npk = (C.c_int64 * n)(*[int(x) for x in npackets])
        st = (MCStats * n)()
        self._chk(self.L.tam_batch_mc_run(self.h, npk, int(seed), int(stream0), int(bool(accumulate)), st))
        return list(st)

    def heat_run(self, loops):
        self._chk(self.L.tam_batch_heat_run(self.h, int(loops)))

    def set_launch(self, tile, kz, blocks_per_cu, rng_batch):
        self._chk(self.L.tam_batch_set_launch(self.h, int(tile), int(kz), int(blocks_per_cu), int(rng_batch)))

    def tune(self, sample_packets, heat_reps=5):
        n = len(self.grids)
        npk = (C.c_int64 * n)(*[int(x) for x in sample_packets])
        t, k, b, r = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        self._chk(self.L.tam_batch_tune(self.h, npk, int(heat_reps), C.byref(t), C.byref(k), C.byref(b),
                                        C.byref(r)))
        return dict(tile=t.value, kz=k.value, blocks_per_cu=b.value, rng_batch=r.value)

    def last_ms(self):
        return self.L.tam_batch_last_ms(self.h)

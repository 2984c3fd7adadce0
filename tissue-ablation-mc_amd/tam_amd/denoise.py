"""ctypes binding of the fluence denoiser (include/tam_denoise.h, in libtam.so).

Feature maps are numpy arrays (4, nz, ny, nx) float32: log(1 + phi_lo),
log10(mu_a), log10(mu_s) - 2, 5 (g - 0.8); targets (nz, ny, nx) float32
log(1 + phi_hi).  The network runs on the GPU only.
"""
import ctypes as C

import numpy as np

from . import lib

EXPORTED = ("tam_dn_create", "tam_dn_destroy", "tam_dn_last_error", "tam_dn_num_params", "tam_dn_get_params",
            "tam_dn_set_params", "tam_dn_store", "tam_dn_store_host", "tam_dn_train", "tam_dn_eval",
            "tam_dn_forward_host", "tam_dn_gradients_host", "tam_dn_apply", "tam_dn_last_ms",
            "tam_dn_predict_slot")

# layer table of the fixed U-Net: (cin, cout, level)
LAYERS = ((16, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2), (96, 32, 1),
          (32, 32, 1), (48, 16, 0), (16, 16, 0))

_bound = False


def _bind():
    global _bound
    L = lib.load()
    if _bound:
        return L
    vp = C.c_void_p
    L.tam_dn_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                C.POINTER(vp)]
    L.tam_dn_destroy.argtypes = [vp]
    L.tam_dn_destroy.restype = None
    L.tam_dn_last_error.argtypes = [vp]
    L.tam_dn_last_error.restype = C.c_char_p
    L.tam_dn_num_params.argtypes = [vp]
    L.tam_dn_num_params.restype = C.c_int64
    L.tam_dn_get_params.argtypes = [vp, vp]
    L.tam_dn_set_params.argtypes = [vp, vp]
    L.tam_dn_store.argtypes = [vp, C.c_int32, vp, C.c_double, vp, C.c_double]
    L.tam_dn_store_host.argtypes = [vp, C.c_int32, vp, vp]
    L.tam_dn_train.argtypes = [vp, C.c_int32, C.c_int32, C.c_double, C.c_uint64, vp]
    L.tam_dn_eval.argtypes = [vp, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.tam_dn_forward_host.argtypes = [vp, vp, vp]
    L.tam_dn_predict_slot.argtypes = [vp, C.c_int32, vp, vp, vp]
    L.tam_dn_gradients_host.argtypes = [vp, vp, vp, vp, C.POINTER(C.c_double)]
    L.tam_dn_apply.argtypes = [vp, vp, C.c_double]
    L.tam_dn_last_ms.argtypes = [vp]
    L.tam_dn_last_ms.restype = C.c_double
    _bound = True
    return L


def param_layout():
    """[(name, offset, shape)] of the FP32 parameter vector."""
    out, off = [], 0
    for i, (ci, co, _) in enumerate(LAYERS):
        out.append((f"w{i}", off, (27, ci, co)))
        off += 27 * ci * co
        out.append((f"b{i}", off, (co,)))
        off += co
    out.append(("head_w", off, (16,)))
    out.append(("head_b", off + 16, (1,)))
    return out


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Denoiser:
    def __init__(self, shape, capacity=0, seed=0, device=0):
        """shape: (nz, ny, nx), each a multiple of 32."""
        self.L = _bind()
        self.shape = tuple(shape)
        nz, ny, nx = self.shape
        h = C.c_void_p()
        rc = self.L.tam_dn_create(device, nx, ny, nz, capacity, seed, C.byref(h))
        if rc != 0:
            raise lib.TamError(rc, self.L.tam_dn_last_error(None).decode())
        self.h = h
        self.n = self.L.tam_dn_num_params(h)

    def _chk(self, rc):
        if rc != 0:
            raise lib.TamError(rc, self.L.tam_dn_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.tam_dn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.empty(self.n, np.float32)
        self._chk(self.L.tam_dn_get_params(self.h, p.ctypes.data))
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.n
        self._chk(self.L.tam_dn_set_params(self.h, p.ctypes.data))

    def store(self, slot, lo, npk_lo, hi, npk_hi):
        """Dataset slot from two TissueGrid transport runs (device to device)."""
        self._chk(self.L.tam_dn_store(self.h, slot, lo.h, float(npk_lo), hi.h, float(npk_hi)))

    def store_host(self, slot, features, target):
        f, t = _f32(features), _f32(target)
        self._chk(self.L.tam_dn_store_host(self.h, slot, f.ctypes.data, t.ctypes.data))

    def train(self, nslots, steps, lr=1e-3, seed=0):
        losses = np.empty(max(steps, 1), np.float32)
        self._chk(self.L.tam_dn_train(self.h, nslots, steps, lr, seed, losses.ctypes.data))
        return losses[:steps]

    def eval(self, slot):
        a, b = C.c_double(), C.c_double()
        self._chk(self.L.tam_dn_eval(self.h, slot, C.byref(a), C.byref(b)))
        return a.value, b.value

    def predict_slot(self, slot):
        """(prediction, target, noisy input), each log(1 + phi) (nz, ny, nx)."""
        a, b, c = (np.empty(self.shape, np.float32) for _ in range(3))
        self._chk(self.L.tam_dn_predict_slot(self.h, slot, a.ctypes.data, b.ctypes.data, c.ctypes.data))
        return a, b, c

    def forward(self, features):
        f = _f32(features)
        out = np.empty(self.shape, np.float32)
        self._chk(self.L.tam_dn_forward_host(self.h, f.ctypes.data, out.ctypes.data))
        return out

    def gradients(self, features, target):
        f, t = _f32(features), _f32(target)
        g = np.empty(self.n, np.float32)
        loss = C.c_double()
        self._chk(self.L.tam_dn_gradients_host(self.h, f.ctypes.data, t.ctypes.data, g.ctypes.data,
                                               C.byref(loss)))
        return g, loss.value

    def apply(self, grid, npk):
        """Denoise a TissueGrid's tally in place (same device and grid)."""
        self._chk(self.L.tam_dn_apply(self.h, grid.h, float(npk)))

    def last_ms(self):
        return self.L.tam_dn_last_ms(self.h)


def features_from(tally, npk, volume_voxel_m3, mua, mus, g):
    """Host-side features of the same definition (for tests and analysis)."""
    phi = tally / (npk * volume_voxel_m3 * 1e6)
    return np.stack([np.log1p(phi), np.log10(np.maximum(mua, 1e-3)), np.log10(np.maximum(mus, 1e-2)) - 2.0,
                     5.0 * (g - 0.8)]).astype(np.float32)

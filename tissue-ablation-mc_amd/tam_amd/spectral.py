"""Spectral sampling helpers of a polychromatic source (ch_opt.f90:28-152)
over the C-ABI (tam_spectrum_*), under the reference's names.

A spectrum is the reference's array(len, 2): here a numpy array of shape
(2, len), row 0 the wavelengths (strictly ascending), row 1 the values --
the same memory as the Fortran column-major table.
"""
import numpy as np

from . import lib


def _table(array):
    a = np.ascontiguousarray(array, dtype=np.float64)
    if a.ndim != 2 or a.shape[0] != 2:
        raise ValueError("spectrum table must have shape (2, len)")
    return a


def _chk(rc):
    if rc != lib.TAM_OK:
        raise lib.TamError(rc, lib.load().tam_last_error(None).decode())


def mk_cdf(array, cdf_last=0.0):
    """mk_cdf (ch_opt.f90:138-152): the normalised trapezoid CDF, cdf[j-1]
    at row j+1; the last entry is cdf_last divided like the others."""
    a = _table(array)
    cdf = np.zeros(a.shape[1])
    cdf[-1] = cdf_last
    _chk(lib.load().tam_spectrum_cdf(lib._ptr(a), a.shape[1], lib._ptr(cdf)))
    return cdf


def sample(array, cdf, n, seed, stream=0, device=0):
    """n draws of sample (ch_opt.f90:28-45) on the GPU, Philox uniforms."""
    a = _table(array)
    c = np.ascontiguousarray(cdf, dtype=np.float64)
    out = np.empty(int(n))
    _chk(lib.load().tam_spectrum_sample(int(device), lib._ptr(a), lib._ptr(c), a.shape[1], int(n), int(seed),
                                        int(stream), lib._ptr(out)))
    return out


def lin_inter_2d(array, x, device=0):
    """search_2D + lin_inter_2D (ch_opt.f90:79-135): the table's value at
    each wavelength of x, on the GPU."""
    a = _table(array)
    xs = np.ascontiguousarray(x, dtype=np.float64).ravel()
    out = np.empty(xs.size)
    _chk(lib.load().tam_spectrum_interp(int(device), lib._ptr(a), a.shape[1], lib._ptr(xs), xs.size,
                                        lib._ptr(out)))
    return out

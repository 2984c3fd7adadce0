"""Write golden vectors from the reference's own Fortran (ran2, the photon
walk, stokes, the thermal fits and the spectral helpers), compiled from the reference sources by
`make -C oracle ref` into oracle/_ref/libref.so.

    python tests/golden/make_reference_fixtures.py

Only data (inputs and outputs) is written, to tests/golden/reference_vectors.json,
so the oracle tests keep their bit-exact pins where the reference sources are
absent.
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_ffi as O  # noqa: E402

SEED = -95648324          # mcpolar.f90:97-98 for rank 0
# an irregular lamp-like spectrum (nm, relative power) and probe values that
# fall below, on and between the CDF's nodes and the wavelength rows
SPEC_X = [400.0, 430.0, 455.0, 500.0, 520.0, 575.0, 600.0, 650.0, 720.0, 800.0, 910.0, 1000.0]
SPEC_Y = [0.1, 0.35, 0.8, 1.0, 0.95, 0.6, 0.7, 1.2, 0.9, 0.4, 0.2, 0.05]
SPEC_VALUES = [0.0, 1e-3, 0.01, 0.05, 0.2, 0.333, 0.5, 0.75, 0.9, 0.99, 0.999999, 1.0, 400.0, 410.0, 455.0,
               517.3, 600.0, 999.0, 1000.0]


def main():
    R = O.ref()
    out = {"source": "reference Fortran compiled by oracle/Makefile target ref (flang -fdefault-real-8)"}
    # ran2 stream
    s = C.c_int(SEED)
    v = np.zeros(64)
    R.ref_ran2_seq(C.byref(s), 64, v)
    out["ran2"] = {"seed": SEED, "values": [float.hex(float(x)) for x in v]}
    # photon walk, compiled 80^3 grid, the input.params extents
    n = C.c_int()
    R.ref_grid_n(C.byref(n))
    n = n.value
    jm = np.zeros(n ** 3)
    esc = C.c_long()
    s = C.c_int(SEED)
    R.ref_mc(20000, 0.03, 0.03, 0.06, C.byref(s), 0, np.zeros((n + 2) ** 3), jm, C.byref(esc))
    jm = jm.reshape(n, n, n)
    out["mc"] = {"n": n, "xmax": 0.03, "ymax": 0.03, "zmax": 0.06, "nphotons": 20000, "seed": SEED,
                 "escapes": int(esc.value),
                 "layer_sums": [float.hex(float(x)) for x in jm.sum(axis=(1, 2))],
                 "sum": float.hex(float(jm.sum())), "sum_sq": float.hex(float((jm ** 2).sum())),
                 "nonzero": int((jm != 0).sum())}
    # stokes: 200 successive scatterings from straight down
    ang = np.array([-1.0, 0.0, 1.0, 0.0, 0.0])
    dirs = np.zeros(3 * 200)
    s = C.c_int(-12345)
    R.ref_stokes(200, 0.9, C.byref(s), ang, dirs)
    out["stokes"] = {"hgg": 0.9, "seed": -12345, "start": [-1.0, 0.0, 1.0, 0.0, 0.0],
                     "dirs": [float.hex(float(x)) for x in dirs]}
    # thermal fits on a grid of inputs
    T = np.linspace(250.0, 600.0, 15)
    W = np.linspace(0.0, 0.75, 15)
    Q = np.linspace(0.0, 2e-6, 15)
    rho = np.linspace(900.0, 1200.0, 15)
    fo = np.zeros(6 * 15)
    R.ref_fits(15, T, W, Q, 1.3e-6, rho, fo)
    out["fits"] = {"T": T.tolist(), "W": W.tolist(), "Q": Q.tolist(), "rho": rho.tolist(), "qvapor": 1.3e-6,
                   "out": [float.hex(float(x)) for x in fo]}
    # spectral helpers (ch_opt.f90:28-152) on an irregular 12-row spectrum
    x, y = SPEC_X, SPEC_Y
    tab = O.spectrum_table(x, y)
    nrow = len(x)
    cdf = np.zeros(nrow)
    cdf[-1] = 0.25                       # the entry mk_cdf leaves to the caller (then divided)
    R.ref_mk_cdf(nrow, tab, cdf)
    waves = np.zeros(200)
    s = C.c_int(-777)
    R.ref_sample(nrow, 200, tab, cdf, C.byref(s), waves)
    vals = np.array(SPEC_VALUES)
    nl = np.zeros(2 * len(vals), np.int32)
    R.ref_search(nrow, tab, cdf, len(vals), vals, nl)
    nl = nl.reshape(2, len(vals))
    # interpolation at values inside each routine's own bracket
    n1 = np.ascontiguousarray(np.clip(nl[0], 1, nrow - 2).astype(np.int32))
    n2 = np.ascontiguousarray(np.clip(nl[1], 1, nrow - 1).astype(np.int32))
    li = np.zeros(2 * len(vals))
    R.ref_lin_inter(nrow, tab, cdf, len(vals), vals, n1, n2, li)
    out["spectral"] = {"x": list(x), "y": list(y), "cdf_last_in": 0.25,
                       "cdf": [float.hex(float(v)) for v in cdf],
                       "sample_seed": -777, "samples": [float.hex(float(v)) for v in waves],
                       "values": list(SPEC_VALUES), "search_1d": nl[0].tolist(), "search_2d": nl[1].tolist(),
                       "lin_inter_1d": [float.hex(float(v)) for v in li[:len(vals)]],
                       "lin_inter_2d": [float.hex(float(v)) for v in li[len(vals):]]}
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "reference_vectors.json"))


if __name__ == "__main__":
    main()

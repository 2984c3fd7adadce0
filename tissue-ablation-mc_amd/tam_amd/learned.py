"""Training data, training schedule and error measures of the fluence
denoiser.

Training tissues (the learned-denoiser configuration): 1-4 planar layers at
random thicknesses; per layer mu_a log-uniform in [0.1, 10] /cm, mu_s
log-uniform in [50, 300] /cm, g uniform in [0.7, 0.95], n = 1.38; air above.
The grid has the voxel size of the coupled 256^3 run (2x2x1 cm / 256^3):
128^3 over 1 x 1 x 0.5 cm.  Half of the tissues carry an ablation crater (a
mu_t = 0 cylinder under the beam), as the coupled run produces.  Gaussian
beam, 1/e^2 radius 0.15 cm.  Inputs come from 1e5 or 2.5e5 packets, targets
from 1e7 (training) or 1e8 (evaluation) packets; each run has its own seed.
"""
import time

import numpy as np

from . import configs, lib
from .denoise import Denoiser

EXT = (1.0, 1.0, 0.5)
FLOPS_FWD = None


def random_layers(rng):
    nl = int(rng.integers(1, 5))
    cuts = np.sort(rng.uniform(0.01, 0.45, nl - 1))
    thick = np.diff(np.concatenate([[0.0], cuts, [1e9]]))
    layers = []
    for t in thick:
        mua = float(np.exp(rng.uniform(np.log(0.1), np.log(10.0))))
        mus = float(np.exp(rng.uniform(np.log(50.0), np.log(300.0))))
        g = float(rng.uniform(0.7, 0.95))
        layers.append((float(t), mua, mus, g, 1.38))
    return layers


def layered_problem(n, layers, crater=None, beam_w=0.15, ext=EXT):
    """Problem on an n^3 grid; crater = (radius_cm, depth_cm) for a cylinder,
    (radius_cm, depth_cm, "bowl") for a paraboloid of that radius and
    central depth, or None."""
    p = configs._config_params((n, n, n), ext)
    p.beam = lib.BEAM_GAUSSIAN
    p.beam_radius = beam_w
    dz = ext[2] / n
    mat_z = configs._layered_mat(n, dz, [l[0] for l in layers])
    mat = np.broadcast_to(mat_z[:, None, None], (n, n, n)).copy()
    table = [l[1:] for l in layers] + [(0.0, 0.0, 0.0, 1.0)]       # last entry: air
    mut = np.array([t[0] + t[1] for t in table])
    rk = np.zeros((n + 2,) * 3)
    rk[1:-1, 1:-1, 1:-1] = mut[mat]
    if crater is not None:
        r, d = crater[:2]
        zz, yy, xx = np.meshgrid((np.arange(n) + 0.5) * dz, (np.arange(n) + 0.5) * ext[1] / n,
                                 (np.arange(n) + 0.5) * ext[0] / n, indexing="ij")
        rho2 = (xx - ext[0] / 2) ** 2 + (yy - ext[1] / 2) ** 2
        depth = d * (1.0 - rho2 / (r * r)) if len(crater) > 2 and crater[2] == "bowl" else d
        inside = (rho2 <= r * r) & (zz >= ext[2] - depth)
        mat[inside] = len(table) - 1
        rk[1:-1, 1:-1, 1:-1][inside] = 0.0
    temp, kappa, density, heatcap = configs._thermal_cube(p, configs.C0_THERMAL)
    return configs.Problem("layered", p, mat, table, temp, kappa, density, heatcap, rhokap=rk, packets=1)


def random_problem(n, rng):
    layers = random_layers(rng)
    crater = None
    if rng.random() < 0.5:
        crater = (float(rng.uniform(0.03, 0.2)), float(rng.uniform(0.02, 0.2)))
    return layered_problem(n, layers, crater), layers, crater


def random_skin_layers(rng):
    """Three-layer skin around the coupled workload's (epidermis 0.01 cm,
    20 / 470 / 0.79; dermis 0.2 cm, 2.7 / 200 / 0.82; subcutis 1.0 / 130 /
    0.8; n = 1.4), each property jittered: the epidermis lies outside the
    1-4-layer draws (mu_a <= 10, mu_s <= 300 /cm)."""
    u = lambda lo, hi: float(rng.uniform(lo, hi))
    lg = lambda lo, hi: float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
    return [(u(0.005, 0.02), lg(10.0, 30.0), lg(300.0, 600.0), u(0.75, 0.85), 1.4),
            (u(0.1, 0.3), lg(1.5, 4.0), lg(150.0, 250.0), u(0.78, 0.86), 1.4),
            (1e9, lg(0.5, 2.0), lg(100.0, 160.0), u(0.75, 0.85), 1.4)]


def random_skin_problem(n, rng):
    layers = random_skin_layers(rng)
    crater = None
    if rng.random() < 0.7:                       # most coupled cycles run on a crater
        crater = (float(rng.uniform(0.03, 0.25)), float(rng.uniform(0.01, 0.2)))
        if rng.random() < 0.5:                   # ablation under a Gaussian beam digs a bowl
            crater = crater + ("bowl",)
    return layered_problem(n, layers, crater), layers, crater


def make_pairs(dn, slots, n, seed, npk_hi, npk_lo_choices=(100_000, 250_000), log=print, kind="layers"):
    """Fill dataset slots with MC pairs; returns per-slot metadata.  kind:
    "layers" (1-4 random planar layers) or "skin" (jittered three-layer skin)."""
    rng = np.random.default_rng(seed)
    meta = []
    draw = random_skin_problem if kind == "skin" else random_problem
    for s in slots:
        t0 = time.perf_counter()
        prob, layers, crater = draw(n, rng)
        npk_lo = int(rng.choice(npk_lo_choices))
        lo, hi = prob.grid(), prob.grid()
        st_lo = lo.mc_run(npk_lo, seed=2 * (seed * 100003 + s) + 1)
        st_hi = hi.mc_run(npk_hi, seed=2 * (seed * 100003 + s) + 2)
        dn.store(s, lo, npk_lo, hi, npk_hi)
        lo.close()
        hi.close()
        meta.append(dict(slot=s, kind=kind, layers=layers, crater=crater, npk_lo=npk_lo, npk_hi=npk_hi,
                         mc_ms=st_hi.kernel_ms + st_lo.kernel_ms, interactions=st_hi.interactions / npk_hi,
                         wall_s=time.perf_counter() - t0))
        log(f"pair {s}: {len(layers)} layers, crater {crater}, lo {npk_lo}, hi {npk_hi:.0e}, "
            f"{st_hi.kernel_ms:.0f} ms, {st_hi.interactions / npk_hi:.0f} interactions/packet")
    return meta


def conv_flops(shape):
    """MFMA FLOP of one forward pass (padded input channels included)."""
    from .denoise import LAYERS
    nz, ny, nx = shape
    f = 0
    for cin, cout, lvl in LAYERS:
        f += 2 * 27 * cin * cout * (nz >> lvl) * (ny >> lvl) * (nx >> lvl)
    return f


def fluence_error(phi, ref, frac=1e-2):
    """Median and 90th percentile of |phi - ref| / ref over voxels with
    ref > frac * max(ref)."""
    sel = ref > frac * ref.max()
    e = np.abs(phi[sel] - ref[sel]) / ref[sel]
    return float(np.median(e)), float(np.percentile(e, 90)), int(sel.sum())


def fluence_bias(phi, ref, frac=1e-2):
    """Signed bias of phi against ref over voxels with ref > frac * max(ref):
    the median of (phi - ref) / ref, and sum(phi) / sum(ref) - 1."""
    sel = ref > frac * ref.max()
    return float(np.median((phi[sel] - ref[sel]) / ref[sel])), float(phi[sel].sum() / ref[sel].sum() - 1.0)

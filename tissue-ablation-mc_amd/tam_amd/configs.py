"""Problem builders: the reference's default run (input.params + the module
defaults of the Fortran) and the benchmark configurations.

Each builder returns a Problem: tam_params, the material grid and table, the
halo temperature/property arrays and the run controls.  The reference-default
builder restates mcpolar.f90:79-146 and initThermalCoeff (3dFD.f90:233-309);
the others describe the physics extensions (scattering, Fresnel, layers,
beam profiles, convective top face) that the reference code lacks.
"""
import dataclasses
import math

import numpy as np

from . import lib

# ---------------------------------------------------------- reference fits --
W0 = 0.75                                   # thermalConst_mod.f90:6
PROTEIN = 1.0 - W0


def air_thermal_cond(T):                    # thermalConst_mod.f90:11-25
    return -0.188521 * np.exp(-0.000367259 * (T - 273.15)) + 0.212453


def air_density(T):                         # :28-38
    return 101.325e3 / (287.058 * T)


def skin_density(W):                        # :54-62
    return 1000.0 / (W + 0.649 * PROTEIN)


def skin_heatcap(W):                        # :65-73
    return 1000.0 * (4.2 * W + 1.09 * PROTEIN)


def skin_cond(W, rho):                      # :76-85
    return rho * (6.28e-4 * W + 1.17e-4 * PROTEIN)


@dataclasses.dataclass
class Problem:
    name: str
    params: lib.Params
    mat: np.ndarray                 # (nz, ny, nx) uint8
    table: list                     # [(mua, mus, g, n)]
    temp: np.ndarray                # halo, K
    kappa: np.ndarray
    density: np.ndarray
    heatcap: np.ndarray
    rhokap: np.ndarray = None       # halo, optional
    packets: int = 0
    steps: int = 0                  # heat steps per cycle
    cycles: int = 1
    power: float = 1.0              # W
    notes: str = ""

    @property
    def scale(self):
        """W/m^3 per unit (mu_a * path length) per packet: P / (N * V)."""
        return self.power / (self.packets * self.params.volume_voxel)

    def grid(self):
        g = lib.TissueGrid(self.params)
        g.set_optics(self.mat, self.table, self.rhokap)
        g.set_thermal(self.temp, self.kappa, self.density, self.heatcap)
        return g


def base_params(nx, ny, nz, xmax, ymax, zmax):
    p = lib.Params()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.xmax, p.ymax, p.zmax = xmax, ymax, zmax
    p.dscale = 1.0
    p.thermal_mode = lib.THERMAL_TABLE
    p.mu_water, p.mu_protein = 510.0, 170.0        # ch_opt.f90:171-172
    p.ablate_K = 1e30
    p.arr_A, p.arr_dE, p.arr_R = 3.1e98, 6.3e5, 8.314   # 3dFD.f90:439-441
    p.arr_Tlo, p.arr_Thi = 43.0 + 273.0, 100.0 + 273.0  # 3dFD.f90:450
    p.arr_thr = (0.53, 1.0, 10000.0)                    # 3dFD.f90:443-445
    p.damage_every_step = 0
    p.beam = lib.BEAM_TOPHAT
    p.beam_radius = 125e-4
    p.n_ext_zlo = 0.0
    p.n_ext_zhi = 0.0
    p.device = 0
    return p


def explicit_limit(alpha, dx, dy, dz, dscale):
    """Largest stable step of u' = dscale * alpha * Laplacian(u) (7-point)."""
    return 1.0 / (2.0 * dscale * alpha * (1.0 / dx ** 2 + 1.0 / dy ** 2 + 1.0 / dz ** 2))


# ----------------------------------------------------- reference default ---

def read_input_params(path):
    """input.params: 14 positional lines (mcpolar.f90:79-94)."""
    with open(path) as f:
        tok = [ln.split()[0] for ln in f if ln.strip()]
    rep = tok[8]
    rep = eval(rep, {"__builtins__": {}}) if "/" in rep else float(rep)   # "10000000./1." in the file
    return dict(nphotons=int(tok[0]), xmax=float(tok[1]), ymax=float(tok[2]), zmax=float(tok[3]),
                n1=float(tok[4]), n2=float(tok[5]), total_time=float(tok[6]), loops=int(tok[7]),
                repetitionRate_1=float(rep), power=float(tok[9]), energyPerPixel=float(tok[10]),
                ablateTemp=float(tok[11]), pulsesToDo=int(tok[12]), pulsetype=tok[13])


REFERENCE_INPUT = dict(nphotons=125000, xmax=0.03, ymax=0.03, zmax=0.06, n1=1.0, n2=1.38,
                       total_time=2.0, loops=1, repetitionRate_1=10000000.0, power=70.0,
                       energyPerPixel=400.0, ablateTemp=500.0, pulsesToDo=1, pulsetype="gaussian")


def reference_default(inp=None, n=80):
    """The reference's own run: mcpolar.f90:61-146 with initThermalCoeff."""
    inp = dict(REFERENCE_INPUT if inp is None else inp)
    xmax, ymax, zmax = inp["xmax"], inp["ymax"], inp["zmax"]
    p = base_params(n, n, n, xmax, ymax, zmax)
    # initThermalCoeff (3dFD.f90:249-291)
    p.dx_m = (2.0 * xmax * 1e-2) / (n + 2.0)
    p.dy_m = (2.0 * ymax * 1e-2) / (n + 2.0)
    p.dz_m = (2.0 * zmax * 1e-2) / (n + 2.0)
    hc = skin_heatcap(W0)
    rho = skin_density(W0)
    k_skin = skin_cond(W0, rho)
    alpha = k_skin / (rho * skin_heatcap(W0))
    constd = 1.0 / p.dx_m ** 2 + 1.0 / p.dy_m ** 2 + 1.0 / p.dz_m ** 2
    p.delt = 1.0 / (1.0 * alpha * constd)
    p.dscale = 0.5                                   # 3dFD.f90:128-129
    p.volume_voxel = (2.0 * xmax * 1e-2 / n) * (2.0 * ymax * 1e-2 / n) * (2.0 * zmax * 1e-2 / n)
    p.mass_voxel = rho * p.volume_voxel
    p.qvapor = 2256e3 * p.mass_voxel
    p.t_ref = 278.0
    p.boil = 1
    p.heat_fp64 = 1                                  # boiling clamp: FP64 temperatures
    p.thermal_mode = lib.THERMAL_WATER
    p.ablate_K = inp["ablateTemp"] + 273.0
    p.beam = lib.BEAM_TOPHAT
    p.beam_radius = 250e-4 / 2.0                     # sourceph.f90:23
    hs = (n + 2, n + 2, n + 2)
    kappa = np.full(hs, air_thermal_cond(25.0 + 273.0))
    kappa[1:-1, 1:-1, 1:-1] = k_skin
    density = np.full(hs, rho)
    heatcap = np.full(hs, hc)
    temp = np.full(hs, 5.0 + 273.0)                  # mcpolar.f90:123-129
    temp[0, :, :] = 25.0 + 273.0                     # bottom face (z = 0)
    temp[-1, :, :] = 25.0 + 273.0                    # top face (z = N+1)
    rhokap = np.zeros(hs)
    rhokap[1:-1, 1:-1, 1:-1] = 510.0 + 170.0         # ch_opt.f90:173, gridset.f90:42
    mat = np.zeros((n, n, n), np.uint8)
    table = [(680.0, 0.0, 0.9, inp["n2"])]
    prob = Problem("reference-default", p, mat, table, temp, kappa, density, heatcap, rhokap,
                   packets=inp["nphotons"], steps=inp["loops"], power=inp["power"])
    prob.inp = inp
    prob.pulse_length = (inp["energyPerPixel"] * 1e-3 * 81.0) / inp["power"]   # 3dFD.f90:287
    return prob


# ------------------------------------------------------- benchmark configs --

C0_THERMAL = dict(k=0.0056 * 100.0, rho=1.07 * 1000.0, c=3.4 * 1000.0)   # W/m/K, kg/m^3, J/kg/K
T_BODY = 37.0 + 273.15
T_AMB = 25.0 + 273.15        # convective top: ambient not given; 25 C (tempAir, thermalConst_mod.f90:5)
H_TOP = 0.001 * 1e4          # 0.001 W/cm^2/K in W/m^2/K


def _layered_mat(nz, dz_cm, layers):
    """Material per z-plane: voxel centre depth below the top face decides the layer."""
    mat_z = np.zeros(nz, np.uint8)
    bounds = np.cumsum([t for t in layers])
    for k in range(nz):
        depth = (nz - k - 0.5) * dz_cm               # k = 0 is the bottom plane
        m = int(np.searchsorted(bounds, depth, side="right"))
        mat_z[k] = min(m, len(layers) - 1)
    return mat_z


def _thermal_cube(p, thermal, t0=T_BODY):
    hs = (p.nz + 2, p.ny + 2, p.nx + 2)
    temp = np.full(hs, t0)
    kappa = np.full(hs, thermal["k"])
    density = np.full(hs, thermal["rho"])
    heatcap = np.full(hs, thermal["c"])
    return temp, kappa, density, heatcap


def _config_params(n, ext_cm, steps_dt_frac=0.9):
    nx, ny, nz = n
    Lx, Ly, Lz = ext_cm
    p = base_params(nx, ny, nz, Lx / 2, Ly / 2, Lz / 2)
    p.dx_m, p.dy_m, p.dz_m = Lx / nx * 1e-2, Ly / ny * 1e-2, Lz / nz * 1e-2
    th = C0_THERMAL
    alpha = th["k"] / (th["rho"] * th["c"])
    p.dscale = 1.0
    p.delt = steps_dt_frac * explicit_limit(alpha, p.dx_m, p.dy_m, p.dz_m, p.dscale)
    p.volume_voxel = p.dx_m * p.dy_m * p.dz_m
    p.mass_voxel = th["rho"] * p.volume_voxel
    p.qvapor = 2256e3 * p.mass_voxel
    p.t_ref = T_BODY
    p.boil = 0
    p.robin_top = 1
    p.h_top = H_TOP
    p.t_amb = T_AMB
    p.thermal_mode = lib.THERMAL_TABLE
    p.removed_inert = 1                              # ablated voxels: no conduction (stability)
    p.n_ext_zhi = 1.0                                # air above
    p.n_ext_zlo = 0.0
    return p


def config0(n=100, packets=1_000_000, steps=1000):
    """Homogeneous 1 cm cube, mua 0.5, mus 100, g 0.9, n 1.38, top-hat R 0.1 cm, 1 W."""
    p = _config_params((n, n, n), (1.0, 1.0, 1.0))
    p.beam = lib.BEAM_TOPHAT
    p.beam_radius = 0.1
    mat = np.zeros((n, n, n), np.uint8)
    table = [(0.5, 100.0, 0.9, 1.38)]
    temp, kappa, density, heatcap = _thermal_cube(p, C0_THERMAL)
    return Problem("c0-homogeneous", p, mat, table, temp, kappa, density, heatcap,
                   packets=packets, steps=steps, cycles=1, power=1.0)


SKIN3 = [(0.01, 20.0, 470.0, 0.79, 1.4),     # epidermis: thickness cm, mua, mus, g, n
         (0.2, 2.7, 200.0, 0.82, 1.4),       # dermis
         (1e9, 1.0, 130.0, 0.8, 1.4)]        # subcutis (rest)


def three_layer_skin(n=(200, 200, 200), ext=(2.0, 2.0, 1.0), packets=100_000_000, steps=10_000,
                     cycles=10, power=1.0, name="c1-three-layer-skin"):
    """Three-layer skin, Gaussian beam 1/e^2 radius 0.15 cm, ablation at 100 C."""
    p = _config_params(n, ext)
    p.beam = lib.BEAM_GAUSSIAN
    p.beam_radius = 0.15
    p.ablate_K = 100.0 + 273.0                       # ablateTemp + 273 (3dFD.f90:334)
    p.arr_A, p.arr_dE = 3.1e98, 6.28e5
    p.damage_every_step = 1
    nx, ny, nz = n
    dz_cm = ext[2] / nz
    mat_z = _layered_mat(nz, dz_cm, [l[0] for l in SKIN3])
    mat = np.broadcast_to(mat_z[:, None, None], (nz, ny, nx)).copy()
    table = [l[1:] for l in SKIN3]
    temp, kappa, density, heatcap = _thermal_cube(p, C0_THERMAL)
    prob = Problem(name, p, mat, table, temp, kappa, density, heatcap,
                   packets=packets, steps=steps, cycles=cycles, power=power)
    prob.notes = (f"epidermis {int((mat_z == 0).sum())} planes of {dz_cm:.5f} cm; "
                  f"dermis {int((mat_z == 1).sum())} planes; power {power} W (not given)")
    return prob


def config1(**kw):
    return three_layer_skin(**kw)


def config4_tissue(**kw):
    kw.setdefault("n", (256, 256, 256))
    kw.setdefault("steps", 5000)
    kw.setdefault("cycles", 20)
    kw.setdefault("name", "c4-three-layer-skin-256")
    return three_layer_skin(**kw)


def homogeneous(n, mua, mus, g, nidx=1.38, ext=1.0, packets=1_000_000, steps=2000, seed_thermal=None):
    """One homogeneous instance (the batched configuration's building block)."""
    p = _config_params((n, n, n), (ext, ext, ext))
    p.beam = lib.BEAM_TOPHAT
    p.beam_radius = 0.1 * ext
    mat = np.zeros((n, n, n), np.uint8)
    table = [(mua, mus, g, nidx)]
    temp, kappa, density, heatcap = _thermal_cube(p, C0_THERMAL)
    return Problem(f"homogeneous-{n}", p, mat, table, temp, kappa, density, heatcap,
                   packets=packets, steps=steps, power=1.0)


def batch_manifest(n=64, seed=20240611):
    """The batched-instances configuration: n homogeneous tissues with grid
    edge from {64, 96, 128, 160}, mu_a log-uniform in [0.1, 10] /cm, mu_s
    log-uniform in [50, 300] /cm, g uniform in [0.7, 0.95], n = 1.38, packets
    log-uniform in [1e6, 1e7]; drawn once from numpy's PCG64 with a fixed seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n):
        out.append(dict(edge=int(rng.choice([64, 96, 128, 160])),
                        mua=float(np.exp(rng.uniform(np.log(0.1), np.log(10.0)))),
                        mus=float(np.exp(rng.uniform(np.log(50.0), np.log(300.0)))),
                        g=float(rng.uniform(0.7, 0.95)),
                        packets=int(round(np.exp(rng.uniform(np.log(1e6), np.log(1e7)))))))
    return out


def batch_problems(manifest, steps=2000):
    """One Problem per manifest entry: a 1 cm cube with config-0 thermal
    constants (convective top, 37 C elsewhere), top-hat beam of radius
    0.1 cm, 1 W."""
    return [homogeneous(m["edge"], m["mua"], m["mus"], m["g"], packets=m["packets"], steps=steps)
            for m in manifest]

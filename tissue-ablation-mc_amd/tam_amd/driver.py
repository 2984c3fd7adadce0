"""The reference's driver (mcpolar.f90) over the C-ABI, in Python.

The same loop as the Fortran host (fortran/mcpolar_gpu.f90): while
time <= total_time, one tam_cycle replay per iteration -- transport of
nphotons packets when the laser is on, normalised by the pulse power
(mcpolar.f90:151-174), `loops` heat steps, Arrhenius, setupThermalCoeff --
then the writer's fields (writer.f90:6-81) and file names.
"""
import math
import os

import numpy as np

from . import configs, lib

SPOTS = 81.0                          # spotsPerRow * spotsPerCol (constants.f90)
SEED = 95648324                       # |iseed| of mcpolar.f90:97


def run_length(inp, pulse_length, delt):
    """Pulse length and end time (3dFD.f90:293-306, mcpolar.f90:133-140)."""
    kind = inp["pulsetype"]
    real_pulse = {"tophat": pulse_length, "gaussian": 20000.0 * pulse_length,
                  "triangular": 2.0 * pulse_length}[kind]
    total = inp["total_time"]
    if kind == "gaussian":
        total = 2.0 * pulse_length * (2.0 * math.sqrt(2.0 * math.log(2.0)))
        real_pulse = total
    elif int(total / delt) <= int(real_pulse / delt):
        total = delt * (real_pulse / delt + 2000.0)
    return real_pulse, total


def step_guard(pulse_length, delt):
    """heat_sim_3D's guard (3dFD.f90:102-105): a pulse shorter than one
    explicit step shortens the step to a hundredth of the pulse.  The
    reference applies it on the first heat call, after the run length was
    set from the original step (mcpolar.f90:133-140); the coefficient
    delt/(rho c) of the first cycle's source still holds the original step
    (initThermalCoeff, 3dFD.f90:282-283) until setupThermalCoeff rebuilds it
    (3dFD.f90:339-345).  Returns the step the heat loop uses."""
    return pulse_length / 100.0 if pulse_length < delt else delt


class PowerFunction:
    """getPwrGaussian / TopHat / Triangular (3dFD.f90:365-421)."""

    def __init__(self, inp, pulse_length):
        self.kind, self.power, self.pl = inp["pulsetype"], inp["power"], pulse_length
        self.ramp_down = False

    def __call__(self, pulse):
        if self.kind == "gaussian":
            fwhm = 2.0 * math.sqrt(2.0 * math.log(2.0))
            mu, sig = fwhm * self.pl, self.pl / fwhm
            return self.power * math.exp(-(pulse.time - mu) ** 2 / (2.0 * sig ** 2))
        if not pulse.laser_flag:
            return 0.0
        if self.kind == "tophat":
            return self.power
        slope = self.power / self.pl
        if not self.ramp_down and pulse.time >= self.pl:
            self.ramp_down = True
        return max(0.0, -slope * pulse.time + 2.0 * self.power) if self.ramp_down else slope * pulse.time


def _text_int(i, w=None):
    s = str(int(i))
    if w is None:
        return s
    return "0" * (w - len(s)) + s if len(s) <= w else s[:w]


def _text_real(x, w):
    return ("%.16f" % x)[:w].strip()


def output_tag(inp, n):
    """The file-name tail of writer.f90:21-25 (utils.f90 str formatting)."""
    return (f"{_text_int(inp['power'])}w-{n}-{_text_real(inp['ablateTemp'], 3)}-"
            f"{_text_int(inp['energyPerPixel'], 3)}-{_text_real(inp['xmax'], 5)}-"
            f"{_text_real(inp['ymax'], 5)}-{_text_real(inp['zmax'], 5)}.dat")


def reference_run(inp=None, n=80, max_cycles=None, seed=SEED):
    """Run the reference-default problem; returns (grid, cycles, pulse)."""
    inp = dict(configs.REFERENCE_INPUT if inp is None else inp)
    prob = configs.reference_default(inp, n)
    p = prob.params
    real_pulse, total = run_length(inp, prob.pulse_length, p.delt)
    delt0 = p.delt
    p.delt = step_guard(prob.pulse_length, delt0)
    g = prob.grid()
    if p.delt != delt0:
        # the first cycle's source coefficient keeps the original step
        coeff = np.zeros_like(prob.density)
        coeff[1:-1, 1:-1, 1:-1] = delt0 / (prob.density * prob.heatcap)[1:-1, 1:-1, 1:-1]
        g.upload(lib.FIELD_COEFF, coeff)
    pulse = lib.Pulse()
    pulse.repetitionRate_1 = inp["repetitionRate_1"]
    pulse.realPulseLength = real_pulse
    pulse.laserOn = 1.0
    pulse.laser_flag = 1
    pulse.pulsesToDo = inp["pulsesToDo"]
    pulse.use_pulse = 1
    power = PowerFunction(inp, prob.pulse_length)
    cycles = 0
    while pulse.time <= total:
        if max_cycles is not None and cycles >= max_cycles:
            break
        npk, scale = 0, 0.0
        if pulse.laser_flag:
            npk = inp["nphotons"]
            scale = (power(pulse) / SPOTS) / (npk * p.volume_voxel)      # mcpolar.f90:174
        g.cycle(npk, seed, cycles, scale, inp["loops"], arr_delt=p.delt, pulse=pulse,
                stats=(cycles + 1) % 256 == 0)
        cycles += 1
    g.synchronize()
    return g, cycles, pulse


def writer_fields(g):
    """The writer's eight arrays (writer.f90:21-79), keyed by file prefix:
    interior fields (nz, ny, nx), temperature with halo in degrees C (T - 273),
    tissue = -1 where rhokap <= 0.1 (mcpolar.f90:190-201)."""
    rk = g.download(lib.FIELD_RHOKAP)[1:-1, 1:-1, 1:-1]
    tissue = g.download(lib.FIELD_TISSUE)
    tissue[rk <= 0.1] = -1.0
    thr = g.download(lib.FIELD_THRESTIME)
    return {"jmean-t": g.download(lib.FIELD_JMEAN), "rhokap-t": rk,
            "temp-t": g.download(lib.FIELD_TEMP) - 273.0, "water-t": g.download(lib.FIELD_WATER),
            "tissue-t": tissue, "/time-t-1-": thr[0], "/time-t-2-": thr[1], "/time-t-3-": thr[2]}


def write_outputs(g, inp, n, outdir):
    """Raw FP64 streams with the reference's names (writer.f90:21-79)."""
    tag = output_tag(inp, n)
    paths = []
    for prefix, a in writer_fields(g).items():
        path = os.path.join(outdir, "") + prefix + tag
        np.ascontiguousarray(a, dtype=np.float64).tofile(path)
        paths.append(path)
    return paths

"""The Fortran drop-in host (ROCm flang, ISO_C_BINDING) on the CPU: its
interoperable types have the C structs' sizes, and its set-up (input.params,
initThermalCoeff's time step, the run length, the writer's file names)
agrees with the Python host."""
import math
import os
import re
import shutil
import subprocess

import pytest

from tam_amd import configs, driver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "tissue-ablation-mc_amd", "fortran")
GOLD = os.path.join(os.path.dirname(__file__), "golden")

needs_host = pytest.mark.skipif(not os.path.exists(os.path.join(FDIR, "mcpolar_gpu")),
                                reason="Fortran host not built (flang missing)")


@needs_host
def test_fortran_types_match_c_structs():
    r = subprocess.run([os.path.join(FDIR, "abi_check")], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0
    assert r.stdout.count("struct ") == 5 and "abi version 1" in r.stdout


@needs_host
def test_driver_setup_matches_python_host(tmp_path):
    res = tmp_path / "res"
    out = tmp_path / "data"
    res.mkdir()
    out.mkdir()
    shutil.copy(os.path.join(GOLD, "input.params"), res / "input.params")
    # zero cycles: reads input.params, sets up, and (with a GPU) writes the
    # initial fields; without one it stops at tam_create with the message
    r = subprocess.run([os.path.join(FDIR, "mcpolar_gpu"), str(res), str(out), "24", "0"],
                       capture_output=True, text=True, timeout=120)
    text = r.stdout + r.stderr
    print(text)
    m = re.search(r"# delt\s+(\S+) total_time\s+(\S+) pulse\s+(\S+)", text)
    assert m, text
    delt, total, pulse = (float(x) for x in m.groups())
    inp = configs.read_input_params(os.path.join(GOLD, "input.params"))
    prob = configs.reference_default(inp, 24)
    assert math.isclose(delt, prob.params.delt, rel_tol=1e-15)
    real_pulse, total_py = driver.run_length(inp, prob.pulse_length, prob.params.delt)
    assert math.isclose(total, total_py, rel_tol=1e-15) and math.isclose(pulse, real_pulse, rel_tol=1e-15)
    tag = re.search(r"# outputs (\S+)", text).group(1)
    assert tag == "jmean-t" + driver.output_tag(inp, 24) == "jmean-t70w-24-500-400-0.030-0.030-0.060.dat"
    if r.returncode == 0:                                  # a GPU was present
        assert len(os.listdir(out)) == 8
    else:
        assert "no HIP device" in text


@needs_host
def test_driver_short_pulse_step_guard(tmp_path):
    """A pulse shorter than the explicit step: the Fortran host prints the
    reference's message and runs with delt = pulse/100, with the run length
    from the original step (3dFD.f90:102-105, mcpolar.f90:133-140)."""
    res = tmp_path / "res"
    res.mkdir()
    lines = open(os.path.join(GOLD, "input.params")).read().splitlines()
    lines[10] = "0.01         energyPerPixel (mJ)"
    lines[13] = "tophat    pulsetype"
    (res / "input.params").write_text("\n".join(lines) + "\n")
    r = subprocess.run([os.path.join(FDIR, "mcpolar_gpu"), str(res), str(tmp_path), "24", "0"],
                       capture_output=True, text=True, timeout=120)
    text = r.stdout + r.stderr
    print(text)
    assert "pulselength smaller than timestep, adjusting..." in text
    delt, total, pulse = (float(x) for x in re.search(r"# delt\s+(\S+) total_time\s+(\S+) pulse\s+(\S+)",
                                                       text).groups())
    inp = configs.read_input_params(str(res / "input.params"))
    prob = configs.reference_default(inp, 24)
    assert math.isclose(delt, driver.step_guard(prob.pulse_length, prob.params.delt), rel_tol=1e-15)
    assert math.isclose(delt, prob.pulse_length / 100.0, rel_tol=1e-15)
    real_pulse, total_py = driver.run_length(inp, prob.pulse_length, prob.params.delt)
    assert math.isclose(total, total_py, rel_tol=1e-15) and math.isclose(pulse, real_pulse, rel_tol=1e-15)


def test_output_tag_follows_the_reference_formatting():
    """str(int, len) zero-pads a short integer and keeps the leading digits of
    a long one; str(real, len) keeps the leading characters of f100.16."""
    inp = dict(configs.REFERENCE_INPUT)
    assert driver.output_tag(inp, 80) == "70w-80-500-400-0.030-0.030-0.060.dat"
    inp.update(power=12.9, energyPerPixel=70.0, ablateTemp=100.5, xmax=0.125, ymax=1.5, zmax=2.0)
    assert driver.output_tag(inp, 200) == "12w-200-100-070-0.125-1.500-2.000.dat"
    inp.update(energyPerPixel=4321.0)
    assert driver.output_tag(inp, 8).split("-")[3] == "432"

"""The C-ABI library on the CPU: it loads, exports every function the
header declares, its struct layouts match the ctypes mirror, argument checks
fail cleanly, and the host-only pulse state machine matches the oracle."""
import ctypes as C
import os
import re

import pytest

import oracle_ffi as O
from tam_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("tam.h", "tam_denoise.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(tam_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_function():
    from tam_amd import denoise
    L = lib.load()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(lib.EXPORTED) | set(denoise.EXPORTED)
    assert sorted(os.listdir(os.path.join(ROOT, "include"))) == ["tam.h", "tam_denoise.h"]


def test_abi_version():
    assert lib.load().tam_abi_version() == 1


@pytest.mark.parametrize("which,cls", [(0, lib.Params), (1, lib.Material), (2, lib.MCStats), (3, lib.Pulse),
                                       (4, lib.Timing)])
def test_struct_layouts_match(which, cls):
    assert lib.load().tam_struct_size(which) == C.sizeof(cls)


def test_bad_arguments_fail_before_touching_a_device():
    L = lib.load()
    p = lib.Params()
    p.nx = 0
    h = C.c_void_p()
    assert L.tam_create(C.byref(p), C.byref(h)) == -1
    assert b"grid dimensions" in L.tam_last_error(None)
    p.nx = p.ny = p.nz = 8
    assert L.tam_create(C.byref(p), C.byref(h)) == -1           # zero extents
    assert L.tam_create(None, C.byref(h)) == -1


def test_create_without_device_reports_it():
    L = lib.load()
    from tam_amd import configs
    p = configs.config0(n=8).params
    h = C.c_void_p()
    rc = L.tam_create(C.byref(p), C.byref(h))
    if rc == 0:
        L.tam_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == -6 and b"no HIP device" in L.tam_last_error(None)


def test_pulse_state_machine_matches_oracle():
    """Multi-pulse top-hat train through tam_pulse_advance (host code of the
    library) and the oracle's restatement of 3dFD.f90:199-214."""
    L = lib.load()
    a = lib.Pulse()
    b = O.Pulse()
    a.realPulseLength = b.realPulseLength = 0.3
    a.repetitionRate_1 = b.repetitionRate_1 = 0.5
    a.pulsesToDo = b.pulsesToDo = 3
    a.laser_flag = b.laser_flag = 1
    a.laserOn = b.laserOn = 1.0
    a.use_pulse = 1
    for step in range(200):
        L.tam_pulse_advance(C.byref(a), 0.05)
        O.lib().or_pulse_advance(C.byref(b), 0.05)
        assert (a.laser_flag, a.pulsesDone, a.laserOn) == (b.laser_flag, b.pulsesDone, b.laserOn), step
        assert (a.time, a.pulseCount, a.repetitionCount) == (b.time, b.pulseCount, b.repetitionCount)
    assert a.pulsesDone == 3

"""The gather-rate probe (libgt_probe.so) that bench.py's gather ceiling
rests on: it times a launch, a larger window is not faster than an
L2-resident one, masking lanes off saves time, and bad arguments are
refused with a message."""
import ctypes as C
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

LIB = Path(__file__).resolve().parent.parent / "graph-tool-rs_amd" / "gt_hip" / "libgt_probe.so"


def _probe():
    lib = C.CDLL(str(LIB))
    lib.gt_probe_gather.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.POINTER(C.c_double), C.POINTER(C.c_uint32)]
    lib.gt_probe_error.restype = C.c_char_p
    return lib


def _time(lib, window, tiles, active):
    us, cus = C.c_double(), C.c_uint32()
    assert lib.gt_probe_gather(window, tiles, active, 3, C.byref(us), C.byref(cus)) == 0, \
        lib.gt_probe_error()
    assert cus.value > 0
    return us.value


def test_probe_times_gathers(gpu):
    lib = _probe()
    resident = _time(lib, 1 << 18, 16384, 1024)
    spread = _time(lib, 1 << 24, 16384, 1024)
    masked = _time(lib, 1 << 24, 16384, 128)
    assert resident > 0 and spread > 0 and masked > 0
    assert spread >= 0.9 * resident
    assert masked < spread


def test_probe_refuses_bad_arguments(gpu):
    lib = _probe()
    us, cus = C.c_double(), C.c_uint32()
    assert lib.gt_probe_gather(1024, 16, 2048, 1, C.byref(us), C.byref(cus)) == -1
    assert b"active" in lib.gt_probe_error()
    assert lib.gt_probe_gather(0, 16, 16, 1, C.byref(us), C.byref(cus)) == -1

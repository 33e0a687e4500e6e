"""Exhaustive GPU check that the traversal's f32 square root and reciprocal
(f64 hardware approximation rounded once to f32) equal the correctly rounded
results for every positive f32 input, plus a CPU-anchored sample."""
import ctypes as C

import numpy as np
import pytest

from galaxy_amd import NbodyGpu
from galaxy_amd import _lib as L

pytestmark = pytest.mark.gpu


def test_sqrt_and_reciprocal_are_correctly_rounded_for_all_f32():
    gpu = NbodyGpu()
    mism = np.zeros(3, np.uint64)
    first = np.zeros(3, np.uint32)
    print("exhaustive rounding check over 2^31 f32 patterns", flush=True)
    rc = L.lib().gbh_check_rounding(gpu._h, mism.ctypes.data_as(C.c_void_p), first.ctypes.data_as(C.c_void_p))
    assert rc == 0, L.lib().gbh_last_error()
    assert mism[0] == 0, f"sqrt mismatches {mism[0]}, first 0x{first[0]:08x}"
    assert mism[1] == 0, f"reciprocal mismatches {mism[1]}, first 0x{first[1]:08x}"
    assert mism[2] == 0, f"normal-range reciprocal mismatches {mism[2]}, first 0x{first[2]:08x}"

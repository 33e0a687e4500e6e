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


def f32_pass(half, theta, d2):
    """The opening criterion of nbody.wgsl:105-108 in IEEE f32 (numpy's float32
    sqrt and division are correctly rounded): fl(2 half / fl(sqrt(d2))) < theta."""
    w = np.float32(2.0) * half.astype(np.float32)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        return (w / np.sqrt(d2.astype(np.float32))) < theta.astype(np.float32)


def test_opening_thresholds_are_the_criterion_boundary():
    """Every per-depth threshold is the first f32 at which the criterion holds:
    it holds at T and not at the f32 just below, for halves across the f32 range
    (the builds' depths 0..16 of many root sizes) and thetas 0.05..2; theta 0
    gives NaN (no d2 passes)."""
    gpu = NbodyGpu()
    rng = np.random.default_rng(7)
    root = np.exp(rng.uniform(np.log(1e-12), np.log(1e17), 3000))
    half = np.concatenate([(root[:, None] * 0.5 ** np.arange(17)).ravel(), [1e-12, 0.5, 1.0, 3.0e18, 1e30]])
    half = half.astype(np.float32)
    theta = np.exp(rng.uniform(np.log(0.05), np.log(2.0), half.size)).astype(np.float32)
    theta[:200] = np.float32(0.5)
    theta[200:400] = np.float32(0.6)
    theta[400:405] = np.float32(1.0)
    theta[405] = np.float32(0.0)
    out = np.zeros(half.size, np.float32)
    print(f"thresholds of {half.size} (half, theta) pairs", flush=True)
    rc = L.lib().gbh_check_thresholds(gpu._h, half.ctypes.data_as(C.c_void_p), theta.ctypes.data_as(C.c_void_p),
                                      half.size, out.ctypes.data_as(C.c_void_p))
    assert rc == 0, L.lib().gbh_last_error()
    none = theta <= 0
    assert np.isnan(out[none]).all()
    t, h, th = out[~none], half[~none], theta[~none]
    # (+inf where no finite d2 passes: 2 half / theta beyond the f32 range)
    assert (t > 0).all() and not np.isnan(t).any()
    assert np.isinf(t[-3:]).any()
    assert f32_pass(h, th, t).all()
    below = np.nextafter(t, np.float32(0))
    first = t == np.float32(np.finfo(np.float32).smallest_subnormal)
    assert not f32_pass(h[~first], th[~first], below[~first]).any()
    gpu.close()

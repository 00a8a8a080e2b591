"""tdlb_cu_reductions_noalloc against tdlb_cu_reductions, through the C ABI.

The no-allocation variant runs the same reduction kernels with their partial
sums in the library's fixed device scratch and writes the result to a device
buffer the caller owns, so its sums must equal tdlb_cu_reductions' bit for bit.
Two calls in a row, the second on a smaller tank, then a third on a 16-bit
handle, check that what the scratch holds from an earlier call never leaks into
a later one.
"""
import ctypes
import sys

import numpy as np
import pytest

import tdlbcpp_amd as td
from tests.helpers import flow_for_tank

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def _reference_sums(cu):
    out = (ctypes.c_double * 8)()
    td.check(td.lib().tdlb_cu_reductions(cu._h, out))
    return np.array(list(out), dtype=np.float64)


def _hip():
    """The HIP runtime the library itself loaded (device buffers for the test)."""
    td.lib()
    with open("/proc/self/maps") as f:
        paths = [ln.split()[-1] for ln in f if "libamdhip64" in ln]
    assert paths, "the library did not load the HIP runtime"
    h = ctypes.CDLL(paths[0])
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


def _noalloc_sums(cu):
    hip = _hip()
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), 64) == 0
    try:
        assert hip.hipMemset(dev, 0xFF, 64) == 0  # NaN in every double until the call writes it
        assert hip.hipDeviceSynchronize() == 0
        td.check(td.lib().tdlb_cu_reductions_noalloc(cu._h, dev))
        cu.synchronize()
        out = np.empty(8, dtype=np.float64)
        assert hip.hipMemcpy(out.ctypes.data, dev, 64, 2) == 0  # hipMemcpyDeviceToHost
        return out
    finally:
        hip.hipFree(dev)


def _tank(T, precision, steps):
    _, gflow = flow_for_tank(T)
    cu = td.ComputeUnit(T, layout=td.LAYOUT_LIJK, streaming=td.STREAM_ESOTWIST, precision=precision, **gflow)
    cu.initialise(1.0)
    cu.bind_rushton(td.rushton_defaults(T))
    cu.run(steps)
    return cu


def test_noalloc_reductions_equal_the_allocating_entry_point():
    runs = [(64, td.PREC_F32, 7), (40, td.PREC_F32, 5), (36, td.PREC_E3M12, 6)]
    for T, prec, steps in runs:
        case(f"reductions_noalloc T={T} precision={prec} steps={steps}")
        cu = _tank(T, prec, steps)
        try:
            got = _noalloc_sums(cu)
            want = _reference_sums(cu)
            np.testing.assert_array_equal(got, want)
            assert np.all(np.isfinite(got)) and got[0] > 0.0 and got[7] != 0.0
            if prec == td.PREC_F32:
                cu.get(td.FIELD_Q)  # the explicit phase reads the stored moments
                np.testing.assert_array_equal(_noalloc_sums(cu), _reference_sums(cu))
        finally:
            cu.close()


def test_noalloc_reductions_reject_a_null_output():
    cu = _tank(36, td.PREC_F32, 1)
    try:
        assert td.lib().tdlb_cu_reductions_noalloc(cu._h, None) == -1  # TDLB_ERR_INVALID
    finally:
        cu.close()

"""The allocation-free entry point (v3dm_find_u8_noalloc, device-resident
inputs and output, working memory in the library's own device array) against
the existing entry point (setLeftFeatures / setRightFeatures /
findPutativeMatches on a context) on the same pairs, bit for bit: two calls
in a row, the second with a smaller input; plus its size limit."""
import ctypes
import sys

import numpy as np
import pytest

import v3dsfmtoolkit_amd as v
from v3dsfmtoolkit_amd._lib import lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu

V3DM_ERR_CAPACITY = -6


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def existing_entry(L, R, rule):
    m = v.SIFTMatching(0)
    try:
        m.enableRoundingWarning(False)
        m.allocate(len(L), len(R))
        m.setKernel(v.KERNEL_I8)
        m.setLeftFeatures(len(L), L)
        m.setRightFeatures(len(R), R)
        pairs = m.findPutativeMatches(*rule)
    finally:
        m.close()
    tab = np.full(len(L) - len(L) % 32, -1, np.int32)
    for k, j in pairs:
        tab[k] = j
    return tab


def _hip():
    """The HIP runtime libv3dmatch.so is linked against (already loaded with
    it), for the test's own device buffers."""
    lib()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


class DeviceBuffer:
    H2D, D2H = 1, 2

    def __init__(self, host: np.ndarray):
        self.h = _hip()
        self.host = np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(self.ptr), max(self.host.nbytes, 16)) == 0
        assert self.h.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, self.H2D) == 0

    def read(self) -> np.ndarray:
        out = np.empty_like(self.host)
        assert self.h.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.ptr)


def noalloc_entry(L, R, rule):
    nL = len(L) - len(L) % 32
    dL, dR = DeviceBuffer(np.asarray(L, np.uint8)), DeviceBuffer(np.asarray(R, np.uint8))
    out = DeviceBuffer(np.full(max(nL, 1), -7, np.int32))  # every entry must be written
    try:
        rc = lib().v3dm_find_u8_noalloc(dL.ptr, len(L), dR.ptr, len(R), rule[0], rule[1], out.ptr, None)
        assert rc == 0, lib().v3dm_last_error(None).decode()
        got = out.read()[:nL]  # hipMemcpy on the null stream waits for the job
    finally:
        for b in (dL, dR, out):
            b.free()
    return got


@pytest.mark.parametrize("rule", [(0.8, 0.8), (0.9, 0.8), (0.8, -1.0)])
def test_noalloc_equals_existing_entry_two_calls(gpu_available, rule):
    big = v.synth_pair(v.synth_seed(1, 0), 8192, 8192, 0.4, 0.05)
    small = v.synth_pair(v.synth_seed(1, 5), 2048, 3000, 0.4, 0.05)  # 3000 -> 2976 rights
    for name, d in [("8192 x 8192", big), ("2048 x 3000", small)]:
        got = noalloc_entry(d["L"], d["R"], rule)
        want = existing_entry(d["L"], d["R"], rule)
        say(f"{name} rule {rule}: {int((got >= 0).sum())} matches")
        np.testing.assert_array_equal(got, want, err_msg=name)
        np.testing.assert_array_equal(got, O.find_u8_blocked(d["L"], d["R"][:len(d["R"]) // 32 * 32], *rule)[2])


def test_noalloc_rejects_more_than_its_array_holds(gpu_available):
    rc = lib().v3dm_find_u8_noalloc(None, 8193, None, 64, 0.8, 0.8, None, None)
    assert rc == V3DM_ERR_CAPACITY
    assert b"at most 8192" in lib().v3dm_last_error(None)

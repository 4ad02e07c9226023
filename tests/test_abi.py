"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, rejects bad handles, and reports a missing device."""
import ctypes
import subprocess

import numpy as np
import pytest

import v3dsfmtoolkit_amd as v
from v3dsfmtoolkit_amd import _lib


def test_library_exports_every_header_symbol():
    names = _lib.header_symbols()
    assert len(names) >= 20
    handle = _lib.lib()
    for n in names:
        assert hasattr(handle, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_bindings_cover_the_header():
    assert set(_lib.header_symbols()) == set(_lib._SIGNATURES)


def test_null_context_is_rejected():
    L = _lib.lib()
    assert L.v3dm_allocate(None, 64, 64, 0) == _lib.V3DM_ERR_ARG
    assert L.v3dm_set_left_u8(None, 0, None) == _lib.V3DM_ERR_ARG
    assert L.v3dm_find_putative_matches(None, 0.9, 0.8, None, 0, None) == _lib.V3DM_ERR_ARG
    assert L.v3dm_find_async(None, 0.9, 0.8) == _lib.V3DM_ERR_ARG
    L.v3dm_destroy(None)  # no-op


def test_create_reports_missing_device_or_succeeds():
    n = v.device_count()
    h = ctypes.c_void_p()
    rc = _lib.lib().v3dm_create(0, ctypes.byref(h))
    if n == 0:
        assert rc == _lib.V3DM_ERR_DEVICE
        assert b"no HIP device" in _lib.lib().v3dm_last_error(None)
        with pytest.raises(v.V3DMError):
            v.SIFTMatching()
    else:
        assert rc == 0
        _lib.lib().v3dm_destroy(h)
    rc = _lib.lib().v3dm_create(-1, ctypes.byref(h))
    assert rc in (_lib.V3DM_ERR_ARG, _lib.V3DM_ERR_DEVICE)


def test_synth_seed_formula():
    assert v.synth_seed(0, 0) == 0x5EED0000
    assert v.synth_seed(3, 5) == 0x5EED0000 + 256 * 3 + 5


def test_synth_rejects_bad_arguments():
    L = _lib.lib()
    assert L.v3dm_synth_pair(1, -1, 10, 0.4, 0.05, 1, None, None, None, None, None) == _lib.V3DM_ERR_ARG
    assert L.v3dm_synth_pair(1, 10, 10, 1.5, 0.05, 1, None, None, None, None, None) == _lib.V3DM_ERR_ARG
    assert L.v3dm_synth_pair(1, 10, 10, 0.4, -1.0, 1, None, None, None, None, None) == _lib.V3DM_ERR_ARG


def test_synth_u8_matches_float_quantisation():
    d = v.synth_pair(77, 64, 64, 0.4, 0.05, f32=True)
    q = np.minimum(255, np.floor(512.0 * d["Lf"].astype(np.float64) + 0.5))
    np.testing.assert_array_equal(d["L"], q.astype(np.uint8))

"""The entry point that allocates nothing (nenya_prove_general_device_noalloc)
returns the same bytes as nenya_prove_general_device, called through the C
ABI twice in a row (the second proof smaller, so it runs over scratch the
first left dirty), and rejects what it cannot hold."""
import ctypes

import numpy as np
import pytest

import nenya_amd as na
from nenya_amd.instances import general_instance

pytestmark = pytest.mark.gpu

_DT = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}


def _noalloc(lib, params, wd, cap):
    out = (ctypes.c_ubyte * cap)()
    n = ctypes.c_size_t()
    rc = lib.nenya_prove_general_device_noalloc(params._h, ctypes.c_void_p(wd.ptr), wd.size, out, cap, ctypes.byref(n))
    return rc, bytes(out[:n.value]) if rc == 0 else n.value


@pytest.mark.parametrize("first,second", [((32, 15), (16, 11)), ((64, 14), (8, 10))])
def test_noalloc_equals_device_entry_point_twice(first, second):
    lib = na._lib._lib()
    for k, log_n in (first, second):
        print(f"noalloc general k={k} log_n={log_n} seed=17", flush=True)
        A, B, Cm, w = general_instance(log_n, k, 17)
        p = na.GeneralSnarkParams.secure(log_n - 5, (A, B, Cm), 0xDEAD, 0xCAFE, k=k)
        wd = na.DeviceArray(w.astype(_DT[k]))
        want = na.prove_general_device(wd.ptr, wd.size, p)
        rc, got = _noalloc(lib, p, wd, len(want) + 4096)
        assert rc == 0, na._lib._lib().nenya_last_error()
        assert got == want
        assert na.verify_general(p, got) == "Ok"


def test_noalloc_rejects_what_it_cannot_hold():
    lib = na._lib._lib()
    print("noalloc limits: log_blocks 11 params, short proof buffer", flush=True)
    A, B, Cm, w = general_instance(16, 32, 18)
    big = na.GeneralSnarkParams.secure(11, (A, B, Cm), 0xDEAD, 0xCAFE, k=32)
    wd = na.DeviceArray(w.astype(np.uint32))
    rc, _ = _noalloc(lib, big, wd, 1 << 20)
    assert rc == 1  # NENYA_E_SHAPE: beyond the scratch's size limit
    A, B, Cm, w = general_instance(10, 32, 19)
    small = na.GeneralSnarkParams.secure(5, (A, B, Cm), 0xDEAD, 0xCAFE, k=32)
    wd = na.DeviceArray(w.astype(np.uint32))
    rc, need = _noalloc(lib, small, wd, 64)
    assert rc == 4 and need > 64  # NENYA_E_ARG, the needed size reported
    rc, got = _noalloc(lib, small, wd, need)
    assert rc == 0 and got == na.prove_general_device(wd.ptr, wd.size, small)

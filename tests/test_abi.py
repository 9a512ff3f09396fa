"""CPU-side checks of the product boundary: the HIP library loads, exports
every symbol include/nenya.h declares, and fails loudly (no CPU fallback)
when no GPU is visible."""
import ctypes
import os
import re

import pytest

import nenya_amd as na

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "nenya.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nenya_\w+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared()
    for n in ("nenya_params_simd_secure", "nenya_prove_simd", "nenya_params_general_secure",
              "nenya_prove_general", "nenya_free", "nenya_params_free", "nenya_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(na.lib_path())
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_fails_loudly_without_gpu():
    if na.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(na.NenyaError) as e:
        na.GeneralSnarkParams.secure(5, (([], [], []),) * 3)
    assert e.value.code == 5


def test_argument_errors_are_codes_not_crashes():
    with pytest.raises(na.NenyaError) as e:
        na.SimdSnarkParams.secure(5, 5, k=12)
    assert e.value.code == 4


def test_mimd_and_verify_argument_errors():
    """Bad widths, a MIMD set without 32 instances and null handles come back
    as NENYA_E_ARG (4), before any device work."""
    one = ([0], [0], [1])
    with pytest.raises(na.NenyaError) as e:
        na.MimdSnarkParams.secure(5, 5, ([one] * 31, [one] * 32, [one] * 32))
    assert e.value.code == 4
    with pytest.raises(na.NenyaError) as e:
        na.MimdSnarkParams.secure(5, 5, ([one] * 32,) * 3, k=12)
    assert e.value.code == 4
    lib = na._lib._lib()
    v = ctypes.c_int(-1)
    assert lib.nenya_verify_general(None, b"", 0, ctypes.byref(v)) == 4
    assert lib.nenya_verify_mimd(None, b"", 0, ctypes.byref(v)) == 4
    assert v.value == -1
    buf, n = ctypes.c_void_p(), ctypes.c_size_t()
    assert lib.nenya_prove_mimd(None, None, ctypes.byref(buf), ctypes.byref(n)) == 4


def test_proof_view_keeps_its_params_alive():
    """A *_view proof reads the params' pinned buffer, so the view holds the
    params object until the view is dropped."""
    import gc
    import weakref

    class Owner:
        pass

    buf = ctypes.create_string_buffer(b"proofbytes")
    owner = Owner()
    ref = weakref.ref(owner)
    v = na._lib._view(ctypes.c_void_p(ctypes.addressof(buf)), ctypes.c_size_t(10), owner)
    del owner
    gc.collect()
    assert ref() is not None and bytes(v) == b"proofbytes" and v[0] == ord("p") and v.readonly
    del v
    gc.collect()
    assert ref() is None


def _container(model, k, s0, s1, seed_w, seed_e, payload, orc):
    import struct
    head = b"NNYA" + struct.pack("<HBBBB6xQQQ", 1, model, k, s0, s1, seed_w, seed_e, len(payload))
    return head + orc.blake3(payload, 32) + payload


def test_proof_container_format(orc):
    """nenya_proof_payload accepts the documented layout (digest from the
    oracle's BLAKE3) and refuses a wrong digest, version or length."""
    payload = bytes(range(256)) * 5
    c = _container(1, 32, 15, 0, 0xDEAD, 0xCAFE, payload, orc)
    got, model, k = na.proof_payload(c)
    assert got == payload and model == 1 and k == 32
    bad = bytearray(c)
    bad[100] ^= 1  # payload byte: digest mismatch
    with pytest.raises(na.NenyaError):
        na.proof_payload(bytes(bad))
    bad = bytearray(c)
    bad[4] = 2  # version
    with pytest.raises(na.NenyaError):
        na.proof_payload(bytes(bad))
    with pytest.raises(na.NenyaError):
        na.proof_payload(c[:-1])  # length

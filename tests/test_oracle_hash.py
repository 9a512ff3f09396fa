"""Pins the oracle's third-party restatements: BLAKE3 (published vectors and
libclang's llvm_blake3_*), the ChaCha block (libsodium chacha20), and the
self-consistency of the rand 0.8.5 stream rules.  The 12-round count, PCG32
seeding, BlockRng buffering and gen_range are parity unpinned: no reference
test or fixture holds an output of them."""
import ctypes
import os

import numpy as np
import pytest

LIBCLANG = "/opt/rocm/lib/llvm/lib/libclang-cpp.so"
LIBSODIUM = "/opt/conda/lib/libsodium.so"


def test_blake3_published_vectors(orc):
    assert orc.blake3(b"").hex() == "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262"
    assert orc.blake3(b"abc").hex() == "6437b3ac38465133ffb63b75273a8db548c558465d79db03fd359c6cd5bd9d85"


@pytest.mark.skipif(not os.path.exists(LIBCLANG), reason="libclang-cpp not present")
def test_blake3_matches_llvm_blake3(orc):
    L = ctypes.CDLL(LIBCLANG)

    class H(ctypes.Structure):
        _fields_ = [("b", ctypes.c_uint8 * 4096)]

    def ref(data, olen, seek=0):
        h = H()
        L.llvm_blake3_hasher_init(ctypes.byref(h))
        L.llvm_blake3_hasher_update(ctypes.byref(h), data, ctypes.c_size_t(len(data)))
        out = (ctypes.c_uint8 * olen)()
        L.llvm_blake3_hasher_finalize_seek(ctypes.byref(h), ctypes.c_uint64(seek), out, ctypes.c_size_t(olen))
        return bytes(out)

    for n in (0, 1, 64, 65, 1023, 1024, 1025, 2048, 3073, 6144, 8192, 16385, 65536 + 7):
        d = bytes(i % 251 for i in range(n))
        assert orc.blake3(d, 200) == ref(d, 200), n


@pytest.mark.skipif(not os.path.exists(LIBSODIUM), reason="libsodium not present")
def test_chacha_block_matches_libsodium():
    from oracle import load
    orc = load()
    S = ctypes.CDLL(LIBSODIUM)
    S.sodium_init()
    key = bytes(range(7, 39))
    out = (ctypes.c_uint8 * 512)()
    S.crypto_stream_chacha20(out, ctypes.c_ulonglong(512), bytes(8), key)
    ref = np.frombuffer(bytes(out), dtype="<u4")
    kw = np.frombuffer(key, dtype="<u4").copy()
    mine = np.concatenate([orc.chacha_block(kw, c, 0, 10) for c in range(8)])
    assert (mine == ref).all()


def test_blockrng_u64_is_two_u32(orc):
    u32s = orc.rng_stream(32, 99, 0, 256)
    u64s = orc.rng_stream(32, 99, 1, 128)
    lo = u32s[0::2]
    hi = u32s[1::2]
    assert (u64s == (lo | (hi << np.uint64(32)))).all()
    # straddle rule: after 63 u32 reads, next_u64 = buf[63] | next_block[0] << 32
    mix = orc.rng_stream(32, 5, 4, 90)
    flat = orc.rng_stream(32, 5, 0, 200)
    pos = 0
    for i, v in enumerate(mix):
        if i % 3 == 0:
            assert v == flat[pos]
            pos += 1
        else:
            assert v == (flat[pos] | (flat[pos + 1] << np.uint64(32)))
            pos += 2


def test_gen_range_bounds_and_permutation(orc):
    v = orc.rng_stream(32, 11, 2, 2000, arg=6)
    assert v.max() <= 6 and len(set(v.tolist())) == 7
    p1, p2, d1, d2 = orc.raa_setup(32, 64, 4, 0xDEAD)
    assert sorted(p1.tolist()) == list(range(256)) and sorted(p2.tolist()) == list(range(256))
    assert (d1 & np.uint64(1)).all() and (d2 & np.uint64(1)).all()


@pytest.mark.parametrize("k", (8, 16, 32, 64))
def test_limb_stream_widths(orc, k):
    lim = orc.rng_stream(k, 3, 3, 64)
    if k < 64:
        u = orc.rng_stream(k, 3, 0, 64)
        assert (lim == (u & np.uint64((1 << k) - 1))).all()
    else:
        assert (lim == orc.rng_stream(k, 3, 1, 64)).all()

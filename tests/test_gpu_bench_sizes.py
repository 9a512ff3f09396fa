"""Byte identity at the sizes the benchmark times: the GPU proofs of the
benchmark's own instances hash to the oracle's digests committed in
tests/golden/oracle_bench.json (General Z/2^32 at 2^20 constraints, SIMD at
n = m = 2^17 per instance for every width, the 2^20-byte PNG batch), and
the 2^25-constraint General proof is accepted by the library's verifier.
Integer work: the bar is bit-exact."""
import hashlib
import json
import os

import numpy as np
import pytest

import nenya_amd as na
from nenya_amd.instances import general_instance, png_instance, simd_instance

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_bench.json")))


def _digest(b):
    return hashlib.blake2b(bytes(b), digest_size=16).hexdigest()


def _limbs(x, k):
    return np.ascontiguousarray(np.asarray(x).astype({8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}[k]))


def test_general_u32_2p20_matches_oracle_digest():
    g = GOLD["general_u32_logn20_seed1"]
    print("general u32 log_n=20 seed=1", flush=True)
    A, B, Cm, w = general_instance(20, 32, 1)
    p = na.GeneralSnarkParams.secure(15, (A, B, Cm), 0xDEAD, 0xCAFE, k=32)
    proof = na.prove_general(w, p)
    assert len(proof) == g["len"] and _digest(proof) == g["digest"]
    assert na.verify_general(p, proof) == "Ok"


@pytest.mark.parametrize("k", (8, 16, 32, 64))
def test_simd_2p17_matches_oracle_digest(k):
    g = GOLD[f"simd_u{k}_log17_seed{301 + k}"]
    print(f"simd k={k} log=17 seed={301 + k}", flush=True)
    A, B, Cm, wit = simd_instance(17, k, 301 + k)
    p = na.SimdSnarkParams.secure(17, 17, 0xDEAD, 0xCAFE, k=k)
    inst = na.SimdR1csInstance(A, B, Cm, wit)
    proof = bytes(na.prove_simd(inst, p, view=True))
    assert len(proof) == g["len"] and _digest(proof) == g["digest"]
    assert na.verify_simd(inst, p, proof) == "Ok"


def test_png_2p20_bytes_matches_oracle_digest():
    g = GOLD["png_u8_logbytes20_gradient"]
    print("png u8 log_bytes=20 gradient", flush=True)
    ln, lm, A, B, Cm, wit = png_instance(20, "gradient")
    p = na.SimdSnarkParams.secure(ln, lm, 0xDEAD, 0xCAFE, k=8)
    inst = na.SimdR1csInstance(A, B, Cm, wit)
    proof = na.prove_simd(inst, p)
    assert len(proof) == g["len"] and _digest(proof) == g["digest"]
    assert na.verify_simd(inst, p, proof) == "Ok"


def test_host_witness_in_chunks_equals_device_witness():
    """A 32 MiB host witness (General Z/2^64 at 2^22 constraints) crosses the
    upload path's 8 MiB chunks through both pinned bounce buffers; the proof
    equals the one made from the same witness already in device memory."""
    print("general u64 log_n=22 seed=5: host witness vs device witness", flush=True)
    A, B, Cm, w = general_instance(22, 64, 5)
    p = na.GeneralSnarkParams.secure(17, (A, B, Cm), 0xDEAD, 0xCAFE, k=64)
    w = _limbs(w, 64)
    host = na.prove_general(w, p)
    wd = na.DeviceArray(w)
    dev = na.prove_general_device(wd.ptr, wd.size, p)
    assert host == dev
    assert na.verify_general(p, host) == "Ok"


def test_general_u32_2p25_verifies():
    print("general u32 log_n=25 seed=1 (verifier)", flush=True)
    A, B, Cm, w = general_instance(25, 32, 1)
    p = na.GeneralSnarkParams.secure(20, (A, B, Cm), 0xDEAD, 0xCAFE, k=32)
    w = _limbs(w, 32)
    del A, B, Cm
    proof = bytes(na.prove_general(w, p, view=True))
    assert na.verify_general(p, proof) == "Ok"
    # one flipped byte in the last opening's columns is rejected
    bad = bytearray(proof)
    bad[len(bad) - 40000] ^= 1
    assert na.verify_general(p, bytes(bad)) != "Ok"

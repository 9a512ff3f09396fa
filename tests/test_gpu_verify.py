"""GPU: the library's verifiers (nenya_verify_*, host checks with the row
recompute on the GPU) accept the GPU's honest proofs and return the same
verdict as the oracle's restated verifier on tampered ones (a flipped byte
in each proof section, a truncated proof)."""
import numpy as np
import pytest

import nenya_amd as na
from nenya_amd.instances import general_instance, mimd_instance, simd_instance

pytestmark = pytest.mark.gpu


def _offsets(n, k):
    """Byte offsets spread over every section of an n-byte proof."""
    eb = 16 * k
    fixed = [5, 40, 70, 100, 128 + 3, 128 + 8 + 2 * eb + 1, 128 + 8 + 7 * eb + 3]
    spread = [int(x) for x in np.linspace(200, n - 1, 24)]
    return sorted({o for o in fixed + spread if 0 <= o < n})


def _flip(p, o):
    b = bytearray(p)
    b[o] ^= 0x01
    return bytes(b)


@pytest.mark.parametrize("k", (32, 8, 64))
def test_verify_simd_matches_oracle(orc, k):
    print(f"verify simd k={k} log=5", flush=True)
    A, B, Cm, wit = simd_instance(5, k, 31)
    params = na.SimdSnarkParams.secure(5, 5, 0xDEAD, 0xCAFE, k=k)
    inst = na.SimdR1csInstance(A, B, Cm, wit)
    p = na.prove_simd(inst, params)
    assert na.verify_simd(inst, params, p) == "Ok"
    assert na.verify_simd(inst, params, p[:-1]) == "Parse"
    for o in _offsets(len(p), k):
        bad = _flip(p, o)
        assert na.verify_simd(inst, params, bad) == orc.simd_verify(k, 5, 5, A, B, Cm, bad), o


@pytest.mark.parametrize("k", (16, 64))
def test_verify_mimd_matches_oracle(orc, k):
    print(f"verify mimd k={k} log=5", flush=True)
    As, Bs, Cs, wit = mimd_instance(5, k, 32)
    params = na.MimdSnarkParams.secure(5, 5, (As, Bs, Cs), 0xDEAD, 0xCAFE, k=k)
    p = na.prove_mimd(wit, params)
    assert na.verify_mimd(params, p) == "Ok"
    assert na.verify_mimd(params, p + b"\0") == "Parse"
    for o in _offsets(len(p), k):
        bad = _flip(p, o)
        assert na.verify_mimd(params, bad) == orc.mimd_verify(k, 5, 5, As, Bs, Cs, bad), o
    w2 = wit.copy()
    w2[4, 3] ^= 1  # instance 4 no longer satisfied
    p2 = na.prove_mimd(w2, params)
    assert na.verify_mimd(params, p2) == orc.mimd_verify(k, 5, 5, As, Bs, Cs, p2) != "Ok"


@pytest.mark.parametrize("k,log_n", [(32, 10), (8, 10), (64, 10)])
def test_verify_general_matches_oracle(orc, k, log_n):
    print(f"verify general k={k} log_n={log_n}", flush=True)
    A, B, Cm, w = general_instance(log_n, k, 33)
    params = na.GeneralSnarkParams.secure(log_n - 5, (A, B, Cm), 0xDEAD, 0xCAFE, k=k)
    p = na.prove_general(w, params)
    assert na.verify_general(params, p) == "Ok"
    assert na.verify_general(params, p[:100]) == "Parse"
    for o in _offsets(len(p), k):
        bad = _flip(p, o)
        assert na.verify_general(params, bad) == orc.general_verify(k, log_n - 5, A, B, Cm, bad), o


def test_verify_general_u32_2p15_blocks_accepts():
    """The headline shape: a 2^20-constraint u32 proof verifies."""
    print("verify general k=32 log_n=20", flush=True)
    A, B, Cm, w = general_instance(20, 32, 1)
    params = na.GeneralSnarkParams.secure(15, (A, B, Cm), 0xDEAD, 0xCAFE, k=32)
    p = na.prove_general(w, params)
    assert na.verify_general(params, p) == "Ok"


@pytest.mark.parametrize("k,log", [(32, 9), (8, 10), (64, 8)])
def test_verify_simd_larger_accepts(orc, k, log):
    """Honest SIMD proofs at sizes where the matrices span many pages (the
    wrapper keeps its converted COO arrays alive through the call)."""
    print(f"verify simd k={k} log={log}", flush=True)
    A, B, Cm, wit = simd_instance(log, k, 34)
    params = na.SimdSnarkParams.secure(log, log, 0xDEAD, 0xCAFE, k=k)
    inst = na.SimdR1csInstance(A, B, Cm, wit)
    p = na.prove_simd(inst, params)
    assert na.verify_simd(inst, params, p) == "Ok"
    bad = bytearray(p)
    bad[len(p) // 2] ^= 1
    assert na.verify_simd(inst, params, bytes(bad)) == orc.simd_verify(k, log, log, A, B, Cm, bytes(bad))


def test_verifiers_accept_containers_of_their_params_only():
    """A proof in the versioned container verifies like the raw bytes; a
    container for other params (seeds), with a changed payload byte, or
    truncated gets verdict Parse."""
    print("containers: general k=32 log_n=10, simd k=16 log=6", flush=True)
    A, B, Cm, w = general_instance(10, 32, 91)
    p = na.GeneralSnarkParams.secure(5, (A, B, Cm), 0xDEAD, 0xCAFE, k=32)
    other = na.GeneralSnarkParams.secure(5, (A, B, Cm), 0xDEAD, 0xBEEF, k=32)
    raw = na.prove_general(w, p)
    c = na.proof_container(p, raw)
    assert len(c) == len(raw) + 72 and na.proof_payload(c) == (raw, 1, 32)
    assert na.verify_general(p, c) == "Ok"
    assert na.verify_general(other, c) == "Parse"
    bad = bytearray(c)
    bad[72 + 5000] ^= 1
    assert na.verify_general(p, bytes(bad)) == "Parse"
    assert na.verify_general(p, c[:-3]) == "Parse"
    A, B, Cm, wit = simd_instance(6, 16, 92)
    ps = na.SimdSnarkParams.secure(6, 6, 0xDEAD, 0xCAFE, k=16)
    inst = na.SimdR1csInstance(A, B, Cm, wit)
    cs = na.proof_container(ps, na.prove_simd(inst, ps))
    assert na.verify_simd(inst, ps, cs) == "Ok"

"""GPU parity: proofs from the HIP prover (through the C ABI) are byte-identical
to the oracle's on the same instance and seeds, and the oracle's restated
verifier accepts them.  Integer work, so the bar is bit-exact."""
import numpy as np
import pytest

import nenya_amd as na
from nenya_amd.instances import general_instance, mimd_instance, png_instance, simd_instance

pytestmark = pytest.mark.gpu


def _first_diff(a, b):
    n = min(len(a), len(b))
    d = np.nonzero(np.frombuffer(a[:n], np.uint8) != np.frombuffer(b[:n], np.uint8))[0]
    return int(d[0]) if len(d) else n


@pytest.mark.parametrize("k", (32, 8, 16, 64))
def test_simd_config0_bytes_equal_oracle(orc, k):
    print(f"simd k={k} log=5 seed=7", flush=True)
    A, B, Cm, wit = simd_instance(5, k, 7)
    params = na.SimdSnarkParams.secure(5, 5, 0xDEAD, 0xCAFE, k=k)
    got = na.prove_simd(na.SimdR1csInstance(A, B, Cm, wit), params)
    want = orc.simd_prove(k, 5, 5, A, B, Cm, wit)
    assert len(got) == len(want) and got == want, f"first differing byte {_first_diff(got, want)}"
    assert orc.simd_verify(k, 5, 5, A, B, Cm, got) == "Ok"


@pytest.mark.parametrize("k,log", [(32, 9), (8, 10), (16, 11), (64, 9), (32, 12)])
def test_simd_larger_bytes_equal_oracle(orc, k, log):
    print(f"simd k={k} log={log} seed=11", flush=True)
    A, B, Cm, wit = simd_instance(log, k, 11)
    params = na.SimdSnarkParams.secure(log, log, 0xDEAD, 0xCAFE, k=k)
    got = na.prove_simd(na.SimdR1csInstance(A, B, Cm, wit), params)
    want = orc.simd_prove(k, log, log, A, B, Cm, wit)
    assert len(got) == len(want) and got == want, f"first differing byte {_first_diff(got, want)}"
    assert orc.simd_verify(k, log, log, A, B, Cm, got) == "Ok"


@pytest.mark.parametrize("k,log_n", [(32, 10), (8, 10), (16, 11), (64, 10), (32, 12)])
def test_general_bytes_equal_oracle(orc, k, log_n):
    print(f"general k={k} log_n={log_n} seed=3", flush=True)
    A, B, Cm, w = general_instance(log_n, k, 3)
    params = na.GeneralSnarkParams.secure(log_n - 5, (A, B, Cm), 0xDEAD, 0xCAFE, k=k)
    got = na.prove_general(w, params)
    want = orc.general_prove(k, log_n - 5, A, B, Cm, w)
    assert len(got) == len(want) and got == want, f"first differing byte {_first_diff(got, want)}"
    assert orc.general_verify(k, log_n - 5, A, B, Cm, got) == "Ok"


@pytest.mark.parametrize("k,log", [(32, 5), (8, 5), (16, 5), (64, 5), (32, 8), (64, 9), (8, 10), (16, 9)])
def test_mimd_bytes_equal_oracle(orc, k, log):
    print(f"mimd k={k} log={log} seed=9", flush=True)
    As, Bs, Cs, wit = mimd_instance(log, k, 9)
    params = na.MimdSnarkParams.secure(log, log, (As, Bs, Cs), 0xDEAD, 0xCAFE, k=k)
    got = na.prove_mimd(wit, params)
    want = orc.mimd_prove(k, log, log, As, Bs, Cs, wit)
    assert len(got) == len(want) and got == want, f"first differing byte {_first_diff(got, want)}"
    assert orc.mimd_verify(k, log, log, As, Bs, Cs, got) == "Ok"


@pytest.mark.parametrize("log_bytes", [14, 16])
def test_png_bytes_equal_oracle(orc, log_bytes):
    print(f"png u8 log_bytes={log_bytes}", flush=True)
    ln, lm, A, B, Cm, wit = png_instance(log_bytes)
    params = na.SimdSnarkParams.secure(ln, lm, 0xDEAD, 0xCAFE, k=8)
    got = na.prove_simd(na.SimdR1csInstance(A, B, Cm, wit), params)
    want = orc.simd_prove(8, ln, lm, A, B, Cm, wit)
    assert len(got) == len(want) and got == want, f"first differing byte {_first_diff(got, want)}"
    assert orc.simd_verify(8, ln, lm, A, B, Cm, got) == "Ok"


def test_proof_views_equal_owned_proofs():
    """The *_view entry points leave the same bytes in the params' pinned
    buffer that the copying entry points return."""
    print("views: simd k=32 log=7, mimd k=16 log=6, general k=8 log_n=10", flush=True)
    A, B, Cm, wit = simd_instance(7, 32, 21)
    ps = na.SimdSnarkParams.secure(7, 7, 0xDEAD, 0xCAFE, k=32)
    inst = na.SimdR1csInstance(A, B, Cm, wit)
    owned = na.prove_simd(inst, ps)
    assert bytes(na.prove_simd(inst, ps, view=True)) == owned
    As, Bs, Cs, mw = mimd_instance(6, 16, 22)
    pm = na.MimdSnarkParams.secure(6, 6, (As, Bs, Cs), 0xDEAD, 0xCAFE, k=16)
    owned = na.prove_mimd(mw, pm)
    assert bytes(na.prove_mimd(mw, pm, view=True)) == owned
    A, B, Cm, w = general_instance(10, 8, 23)
    pg = na.GeneralSnarkParams.secure(5, (A, B, Cm), 0xDEAD, 0xCAFE, k=8)
    owned = na.prove_general(w, pg)
    assert bytes(na.prove_general(w, pg, view=True)) == owned


def test_proofs_match_committed_golden_digests():
    """The GPU proofs of the committed fixtures hash to the digests stored in
    tests/golden/oracle_c0.json (no oracle call at run time)."""
    import hashlib
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_c0.json")))

    def dg(b):
        return hashlib.blake2b(b, digest_size=16).hexdigest()

    print("golden simd k=32 log=5 seed=7", flush=True)
    A, B, Cm, wit = simd_instance(5, 32, 7)
    p = na.prove_simd(na.SimdR1csInstance(A, B, Cm, wit), na.SimdSnarkParams.secure(5, 5, 0xDEAD, 0xCAFE, k=32))
    assert len(p) == g["simd_u32_log5_seed7"]["len"] and dg(p) == g["simd_u32_log5_seed7"]["digest"]
    print("golden general k=32 log_n=10 seed=3", flush=True)
    A, B, Cm, w = general_instance(10, 32, 3)
    p = na.prove_general(w, na.GeneralSnarkParams.secure(5, (A, B, Cm), 0xDEAD, 0xCAFE, k=32))
    assert dg(p) == g["general_u32_logn10_seed3"]["digest"]
    print("golden mimd k=64 log=5 seed=21", flush=True)
    As, Bs, Cs, wit = mimd_instance(5, 64, 21)
    p = na.prove_mimd(wit, na.MimdSnarkParams.secure(5, 5, (As, Bs, Cs), 0xDEAD, 0xCAFE, k=64))
    assert len(p) == g["mimd_u64_log5_seed21"]["len"] and dg(p) == g["mimd_u64_log5_seed21"]["digest"]

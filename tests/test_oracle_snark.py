"""End-to-end oracle checks: SIMD and General proofs are accepted by the
restated verifier, tampering is rejected with the reference's error variant
(snark-ring/tests/end_to_end_{simd,general}.rs), proofs are deterministic,
and the oracle still reproduces its committed golden digests."""
import hashlib
import json
import os

import pytest

from nenya_amd.instances import general_instance, simd_instance

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_c0.json")


def digest(b):
    return hashlib.blake2b(b, digest_size=16).hexdigest()


@pytest.mark.parametrize("k", (8, 16, 32, 64))
def test_simd_accept_and_reject(orc, k):
    A, B, C, wit = simd_instance(5, k, 11)
    p = orc.simd_prove(k, 5, 5, A, B, C, wit)
    assert orc.simd_verify(k, 5, 5, A, B, C, p) == "Ok"
    assert p == orc.simd_prove(k, 5, 5, A, B, C, wit)
    bad = bytearray(p)
    bad[64 + 10] ^= 1  # inside mult_proof.claimed_sum
    assert orc.simd_verify(k, 5, 5, A, B, C, bytes(bad)) != "Ok"
    w2 = wit.copy()
    w2[3, 1] ^= 1  # an input of instance 3 no longer matches its aux variables
    p2 = orc.simd_prove(k, 5, 5, A, B, C, w2)
    assert orc.simd_verify(k, 5, 5, A, B, C, p2) == "MultSumcheck"


@pytest.mark.parametrize("k", (32, 8))
def test_general_accept_and_reject(orc, k):
    A, B, C, w = general_instance(10, k, 5)
    p = orc.general_prove(k, 5, A, B, C, w)
    assert orc.general_verify(k, 5, A, B, C, p) == "Ok"
    bad = bytearray(p)
    bad[4 * 32 + 8 + 3] ^= 1  # first round message of the mult check
    assert orc.general_verify(k, 5, A, B, C, bytes(bad)) == "MultSumcheck"
    w2 = w.copy()
    w2[2] ^= 1
    p2 = orc.general_prove(k, 5, A, B, C, w2)
    assert orc.general_verify(k, 5, A, B, C, p2) == "KernelCheck"


def test_golden_digests(orc):
    g = json.load(open(GOLDEN))
    A, B, C, wit = simd_instance(5, 32, 7)
    p = orc.simd_prove(32, 5, 5, A, B, C, wit)
    assert len(p) == g["simd_u32_log5_seed7"]["len"] and digest(p) == g["simd_u32_log5_seed7"]["digest"]
    A, B, C, w = general_instance(10, 32, 3)
    p = orc.general_prove(32, 5, A, B, C, w)
    assert digest(p) == g["general_u32_logn10_seed3"]["digest"]
    assert [int(x) for x in orc.rng_stream(32, 0, 0, 8)] == g["rng_u32_seed0"]
    from nenya_amd.instances import mimd_instance
    As, Bs, Cs, wit = mimd_instance(5, 64, 21)
    p = orc.mimd_prove(64, 5, 5, As, Bs, Cs, wit)
    assert len(p) == g["mimd_u64_log5_seed21"]["len"] and digest(p) == g["mimd_u64_log5_seed21"]["digest"]


@pytest.mark.parametrize("k", (8, 16, 32, 64))
def test_mimd_accept_and_reject(orc, k):
    """snark-ring/tests/end_to_end_mimd.rs: honest proofs verify; a changed
    final_a, root_estack or stacked-error m-bar is rejected."""
    from nenya_amd.instances import mimd_instance
    As, Bs, Cs, wit = mimd_instance(5, k, 21)
    p = orc.mimd_prove(k, 5, 5, As, Bs, Cs, wit)
    assert orc.mimd_verify(k, 5, 5, As, Bs, Cs, p) == "Ok"
    assert p == orc.mimd_prove(k, 5, 5, As, Bs, Cs, wit)
    nb = k // 8
    eb = 128 * nb
    fin_a = 128 + 8 + 3 * 5 * eb  # mult_proof.final_a
    bad = bytearray(p)
    bad[fin_a] ^= 1
    assert orc.mimd_verify(k, 5, 5, As, Bs, Cs, bytes(bad)) != "Ok"
    bad = bytearray(p)
    bad[96] ^= 1  # root_estack
    assert orc.mimd_verify(k, 5, 5, As, Bs, Cs, bytes(bad)) != "Ok"
    # open_estack is the last section: 8-byte k2 then its m-bar
    lin = eb + 8 + 3 * 5 * eb + 2 * eb
    head = 128 + (8 + 3 * 5 * eb + 4 * eb) + lin
    bad = bytearray(p)
    est = len(p) - _estack_len(orc, k, 5, p, head)
    bad[est + 8] ^= 1
    assert orc.mimd_verify(k, 5, 5, As, Bs, Cs, bytes(bad)) != "Ok"


def _estack_len(orc, k, log_n, p, head):
    """Length of the final (stacked error) opening, read from its headers."""
    import numpy as np
    nb = k // 8
    off = head
    lens = []
    for col_elem in (32, 128, 32, 128):  # open_w, open_emult, open_abc, open_estack
        start = off
        k2 = int(np.frombuffer(p[off:off + 8], np.uint64)[0])
        off += 8 + k2 * 128 * nb
        t = int(np.frombuffer(p[off:off + 8], np.uint64)[0])
        off += 8 + 4 * t + 8
        for _ in range(t):
            cl = int(np.frombuffer(p[off:off + 8], np.uint64)[0])
            off += 8 + cl * col_elem * nb
        off += 8
        for _ in range(t):
            pl = int(np.frombuffer(p[off:off + 8], np.uint64)[0])
            off += 8 + 32 * pl
        lens.append(off - start)
    assert off == len(p)
    return lens[-1]


@pytest.mark.parametrize("k", (8, 32))
def test_mimd_unsatisfied_is_rejected(orc, k):
    """A witness that breaks one instance's constraints makes psi(e_mult)
    nonzero, which the kernel check on the opened columns catches."""
    from nenya_amd.instances import mimd_instance
    As, Bs, Cs, wit = mimd_instance(5, k, 22)
    w2 = wit.copy()
    w2[7, 2] ^= 1
    p = orc.mimd_prove(k, 5, 5, As, Bs, Cs, w2)
    assert orc.mimd_verify(k, 5, 5, As, Bs, Cs, p) != "Ok"


@pytest.mark.parametrize("k", (16, 64))
def test_mimd_packed_relation_identity(orc, k):
    """The MIMD reduction (reduce.rs:143-230) from ring primitives alone:
    for the union-support packing A_hat[r,c] = pack_r2(A_0[r,c]..A_31[r,c]),
    sum_c A_hat[r,c] w_phi[c] = pack_r2((A_k w_k)[r]) + q1 v_alpha + q2 v_beta
    with (q1, q2) the summed mul_with_quotients_only terms, and
    psi(A_hat w_phi) = (A_k w_k)[r]."""
    import numpy as np
    from nenya_amd.instances import mimd_instance, _matvec
    n = 32
    As, Bs, Cs, wit = mimd_instance(5, k, 23)
    cells = {}
    for s in range(32):
        r, c, v = As[s]
        for i in range(len(r)):
            key = (int(r[i]), int(c[i]))
            cells.setdefault(key, np.zeros(32, np.uint64))
            cells[key][s] = (int(cells[key][s]) + int(v[i])) % (1 << k)
    wslots = np.ascontiguousarray(wit.T)
    wphi = orc.pack(k, wslots)
    va = np.zeros(128, np.uint64)
    va[:4] = orc.consts(k, 2)
    vb = orc.consts(k, 1)
    aw = np.stack([_matvec(As[s], wit[s], n, k) for s in range(32)], axis=1)
    lhs = np.zeros((n, 128), np.uint64)
    q1 = np.zeros((n, 128), np.uint64)
    q2 = np.zeros((n, 128), np.uint64)
    msk = (1 << k) - 1
    for (r, c), vals in sorted(cells.items()):
        ahat = orc.pack(k, vals.reshape(1, 32))[0]
        lhs[r] = (lhs[r] + orc.gr_mul(k, ahat, wphi[c])) & np.uint64(msk)
        d1, d2 = orc.quotients(k, vals, wslots[c])
        q1[r] = (q1[r] + d1) & np.uint64(msk)
        q2[r] = (q2[r] + d2) & np.uint64(msk)
    for r in range(n):
        rhs = orc.pack(k, aw[r].reshape(1, 32))[0]
        rhs = (rhs + orc.gr_mul(k, q1[r], va) + orc.gr_mul(k, q2[r], vb)) & np.uint64(msk)
        assert np.array_equal(lhs[r], rhs), r
        assert np.array_equal(orc.psi(k, lhs[r].reshape(1, 128))[0], aw[r])


def test_png_instance(orc):
    """The PNG-filter batch (png_filter.rs): raw bytes equal the example's
    own generator (lines 67-73, restated here as scalar code), every image
    satisfies its Sub/Up recurrence circuit, and the oracle's SIMD proof of
    the batch verifies."""
    from nenya_amd.instances import is_satisfied, png_instance
    M = (1 << 64) - 1

    def rnd_byte(seed):
        x = (seed + 0x9E3779B97F4A7C15) & M
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M
        return (x ^ (x >> 31)) & 0xFF

    ln, lm, A, B, C, w = png_instance(14)
    assert (ln, lm) == (9, 11)  # b = 512 bytes per image, m = 2^ceil_log2(1 + 2b)
    b = 1 << ln
    for s in (0, 7, 31):
        for i in (0, 1, 255, 256, 300, b - 1):
            assert int(w[s, 1 + 2 * i]) == rnd_byte(((s << 40) ^ (i * 0x10000001)) & M)
        # row 1 (odd) uses Up: Filt = Orig - Orig[above]
        i = 256 + 3
        assert int(w[s, 2 + 2 * i]) == (int(w[s, 1 + 2 * i]) - int(w[s, 1 + 2 * (i - 256)])) % 256
        assert is_satisfied(A, B, C, w[s], b, 8)
    p = orc.simd_prove(8, ln, lm, A, B, C, w)
    assert orc.simd_verify(8, ln, lm, A, B, C, p) == "Ok"

"""Pins the oracle's PCS layer (pcs-ring + r2_pcs): query size KAT, shape
selection at the configs' sizes, RAA linearity and commutation with pack and
slot rotation, opening value == direct MLE evaluation, stacked == weighted
sum, rotated opening == real commitment to the rotated message, and tamper
rejection."""
import numpy as np
import pytest

from conftest import WIDTHS, rand_limbs


def test_query_size_kat(orc):
    assert orc.query_size(0.19, 100) == 1060


def test_shapes_at_config_sizes(orc):
    # log 15 / log 17 at u32 (pcs.rs:93-162, r2_pcs.rs:57-62)
    assert orc.pcs_shape(32, 15, 1, 32)[:2] == (64, 512)
    assert orc.pcs_shape(32, 15, 2, 3)[:2] == (8, 4096)
    assert orc.pcs_shape(32, 15, 3)[:2] == (4, 8192)
    assert orc.pcs_shape(32, 17, 0)[:2] == (16, 8192)
    assert orc.pcs_shape(32, 17, 3)[:2] == (8, 16384)
    assert orc.pcs_shape(32, 15, 0) == (8, 4096, 1060)


def mle_eval(orc, k, tab, z):
    eq = orc.eq_table(k, z)
    m = (1 << k) - 1
    acc = sum(orc.gr_mul(k, eq[h], tab[h]).astype(object) for h in range(len(tab)))
    return np.array([int(v) & m for v in acc], np.uint64)


@pytest.mark.parametrize("k", WIDTHS)
def test_raa_linear_and_commutes(orc, rng, k):
    kk, q, seed = 16, 4, 77
    m = np.uint64((1 << k) - 1) if k < 64 else np.uint64(2**64 - 1)
    a = rand_limbs(rng, k, (kk, 32))
    b = rand_limbs(rng, k, (kk, 32))
    with np.errstate(over="ignore"):
        ea, eb, eab = (orc.raa_encode(k, kk, q, seed, x, True) for x in (a, b, (a + b) & m))
        assert (eab == ((ea + eb) & m)).all()
    # pack(encode_packed(a)) == encode(pack(a))
    lhs = orc.pack(k, ea)
    rhs = orc.raa_encode(k, kk, q, seed, orc.pack(k, a), False)
    assert (lhs == rhs).all()
    # rotation commutes
    rot = np.roll(a, -3, axis=1)
    assert (orc.raa_encode(k, kk, q, seed, rot, True) == np.roll(ea, -3, axis=1)).all()


@pytest.mark.parametrize("k", WIDTHS)
@pytest.mark.parametrize("mode", [0, 2])
def test_open_value_is_mle_eval(orc, rng, k, mode):
    k1, k2 = 4, 8
    packed = mode == 0
    msg = rand_limbs(rng, k, (k1 * k2, 32 if packed else 128))
    z = rand_limbs(rng, k, (5, 128))
    ok, vals, _ = orc.pcs_roundtrip(k, k1, k2, 12, 9, mode, 1, msg, z)
    assert ok
    tab = orc.pack(k, msg) if packed else msg
    assert (vals[0] == mle_eval(orc, k, tab, z)).all()
    bad, _, _ = orc.pcs_roundtrip(k, k1, k2, 12, 9, mode, 1, msg, z, tamper=True)
    assert not bad


@pytest.mark.parametrize("k", (8, 32))
@pytest.mark.parametrize("mode", [1, 3])
def test_stacked_is_weighted_sum(orc, rng, k, mode):
    k1, k2, s = 2, 8, 3
    packed = mode == 1
    msgs = rand_limbs(rng, k, (s, k1 * k2, 32 if packed else 128))
    z = rand_limbs(rng, k, (4, 128))
    w = rand_limbs(rng, k, (s, 128))
    ok, vals, _ = orc.pcs_roundtrip(k, k1, k2, 10, 5, mode, s, msgs, z, w)
    assert ok
    m = (1 << k) - 1
    acc = 0
    for t in range(s):
        tab = orc.pack(k, msgs[t]) if packed else msgs[t]
        acc = acc + orc.gr_mul(k, w[t], mle_eval(orc, k, tab, z)).astype(object)
    assert (vals[0] == np.array([int(v) & m for v in acc], np.uint64)).all()


@pytest.mark.parametrize("k", (16, 64))
def test_rotated_open_equals_rotated_message(orc, rng, k):
    k1, k2, nl = 2, 8, 4
    msg = rand_limbs(rng, k, (k1 * k2, 32))
    z = rand_limbs(rng, k, (4, 128))
    ok, ys, _ = orc.pcs_roundtrip(k, k1, k2, 10, 3, 4, nl, msg, z)
    assert ok
    for l in range(nl):
        tab = orc.pack(k, np.roll(msg, -l, axis=1))
        assert (ys[l] == mle_eval(orc, k, tab, z)).all()
    bad, _, _ = orc.pcs_roundtrip(k, k1, k2, 10, 3, 4, nl, msg, z, tamper=True)
    assert not bad


def test_proof_size_formula(orc):
    # m-bar k2 x 512 B, 1060 columns of k1 x 128 B, paths of log2(n) hashes
    assert orc.proof_size_bytes_for(8, 4096, 16384, 1060, 512, 128) == \
        32 + 4096 * 512 + 1060 * (8 + 8 * 128) + 1060 * (8 + 14 * 32) + 1060 * 4

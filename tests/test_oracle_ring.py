"""Pins the oracle's ring layer (galois-ring + reduction constants/embedding).

Checks come from the reference's own known answers (T1 table, 15^-1) and
from what the mathematics fixes: an independent Horner-in-beta product that
uses only the defining relations X^4 = -X - 1 and Y^32 = -Y^3 - Y - alpha,
ring laws, Lagrange/quotient identities and psi(phi(a)) = a.
"""
import numpy as np
import pytest

from conftest import WIDTHS, rand_limbs

# teichmuller.rs:21-38 -- the reference's hardcoded u32 Teichmuller set.
T1_U32 = [
    [0x00000000, 0x00000000, 0x00000000, 0x00000000],
    [0x00000001, 0x00000000, 0x00000000, 0x00000000],
    [0x332608c0, 0x65f5b943, 0x9eff0726, 0xae851724],
    [0x8b615184, 0xcfa3ff54, 0x1296acc9, 0xced422d4],
    [0xd89a58de, 0xafdf9bae, 0xce5fd6c8, 0x20cdcbd3],
    [0x7940e31f, 0x6742fcc7, 0x5a734460, 0xb6a8e4f8],
    [0xf73af71c, 0x50bc4ae9, 0x9e74517f, 0x49a3f426],
    [0x4d21cf70, 0x2674d304, 0x372e9f8b, 0x118269eb],
    [0xacaf044d, 0x96867b2f, 0x6758e472, 0xd141a4ed],
    [0x8940a031, 0x63234aa2, 0xf3f707b1, 0xcbfde110],
    [0x12542d82, 0x8fd60bd1, 0x74611ae8, 0x6dc59203],
    [0x08c508e3, 0xaf43b517, 0x618bae81, 0xb65c0bda],
    [0x7fb827d0, 0x739573b9, 0x29d2594f, 0x3ff87ef1],
    [0xc7efaa2f, 0x99d5857d, 0x86106ec5, 0x5fea383f],
    [0x9e2506f1, 0xecfd19b0, 0x3da92187, 0x1334531d],
    [0x746aef5f, 0x08e6f768, 0x312ba0b8, 0xdb918905],
]


# ---------------- independent pure-Python arithmetic ----------------
def g4mul_py(a, b, k):
    m = (1 << k) - 1
    c = [0] * 7
    for i in range(4):
        for j in range(4):
            c[i + j] += int(a[i]) * int(b[j])
    # reduce from the top with X^4 = -X - 1
    for d in (6, 5, 4):
        h = c[d]
        c[d] = 0
        c[d - 4] -= h
        c[d - 3] -= h
    return [x & m for x in c[:4]]


def grmul_py(a, b, k):
    """Horner in beta: acc = acc*beta + a*b_j, multiplication by beta via
    Y^32 = -Y^3 - Y - alpha.  Independent of the oracle's schoolbook."""
    m = (1 << k) - 1
    A = [[int(x) for x in a[4 * i:4 * i + 4]] for i in range(32)]
    acc = [[0] * 4 for _ in range(32)]
    alpha = [0, 1, 0, 0]
    for j in range(31, -1, -1):
        top = acc[31]
        acc = [[0] * 4] + acc[:31]
        at = g4mul_py(top, alpha, k)
        for c in range(4):
            acc[0][c] = (acc[0][c] - at[c]) & m
            acc[1][c] = (acc[1][c] - top[c]) & m
            acc[3][c] = (acc[3][c] - top[c]) & m
        bj = [int(x) for x in b[4 * j:4 * j + 4]]
        for i in range(32):
            p = g4mul_py(A[i], bj, k)
            acc[i] = [(acc[i][c] + p[c]) & m for c in range(4)]
    return np.array([x for s in acc for x in s], np.uint64)


def one():
    x = np.zeros(128, np.uint64)
    x[0] = 1
    return x


def beta_pow(e, orc, k):
    x = one()
    b = np.zeros(128, np.uint64)
    b[4] = 1
    for _ in range(e):
        x = orc.gr_mul(k, x, b)
    return x


# ---------------- known answers ----------------
def test_t1_u32_matches_reference_table(orc):
    assert orc.t1(32).tolist() == T1_U32


@pytest.mark.parametrize("k", WIDTHS)
def test_t1_fixed_by_x16_and_distinct_mod2(orc, k):
    t1 = orc.t1(k)
    for t in t1:
        x = t.copy()
        for _ in range(4):
            x = orc.g4_mul(k, x, x)
        assert (x == t).all()
    # reductions mod 2 are the 16 distinct elements of F_16
    assert len({tuple(int(v) & 1 for v in t) for t in t1}) == 16


def test_inv15_kats(orc):
    assert orc.inv_2adic(32, 15) == 0xEEEEEEEF
    assert orc.inv_2adic(8, 15) == 239


@pytest.mark.parametrize("k", (8, 16))
def test_inv_2adic_exhaustive(orc, k):
    m = (1 << k) - 1
    step = 1 if k == 8 else 37
    for x in range(1, 1 << k, 2 * step):
        assert (x * orc.inv_2adic(k, x)) & m == 1


# ---------------- products ----------------
@pytest.mark.parametrize("k", WIDTHS)
def test_gr_mul_matches_independent_horner(orc, rng, k):
    for _ in range(2):
        a = rand_limbs(rng, k, 128)
        b = rand_limbs(rng, k, 128)
        assert (orc.gr_mul(k, a, b) == grmul_py(a, b, k)).all()


@pytest.mark.parametrize("k", WIDTHS)
def test_defining_relations(orc, k):
    m = (1 << k) - 1
    # X^4 = -X - 1 in GR(2^k,4)
    x = np.array([0, 1, 0, 0], np.uint64)
    x3 = orc.g4_mul(k, orc.g4_mul(k, x, x), x)
    assert orc.g4_mul(k, x3, x).tolist() == [m, m, 0, 0]
    # f2(beta) = beta^32 + beta^3 + beta + alpha = 0
    b32 = beta_pow(32, orc, k)
    f = b32.astype(object) + beta_pow(3, orc, k).astype(object) + beta_pow(1, orc, k).astype(object)
    f[1] += 1
    assert all(int(v) & m == 0 for v in f)


@pytest.mark.parametrize("k", WIDTHS)
def test_ring_laws(orc, rng, k):
    a, b, c = (rand_limbs(rng, k, 128) for _ in range(3))
    m = np.uint64((1 << k) - 1) if k < 64 else np.uint64(2**64 - 1)
    ab = orc.gr_mul(k, a, b)
    assert (ab == orc.gr_mul(k, b, a)).all()
    assert (orc.gr_mul(k, ab, c) == orc.gr_mul(k, a, orc.gr_mul(k, b, c))).all()
    with np.errstate(over="ignore"):
        bc = (b + c) & m
        lhs = orc.gr_mul(k, a, bc)
        rhs = (ab + orc.gr_mul(k, a, c)) & m
    assert (lhs == rhs).all()
    assert (orc.gr_mul(k, a, one()) == a).all()


@pytest.mark.parametrize("k", WIDTHS)
def test_mul_by_gr1_gr4_equal_lifted_product(orc, rng, k):
    a = rand_limbs(rng, k, 128)
    s = rand_limbs(rng, k, 4)
    lift1 = np.zeros(128, np.uint64)
    lift1[0] = s[0]
    assert (orc.gr_mul_gr1(k, a, s[0]) == orc.gr_mul(k, a, lift1)).all()
    lift4 = np.zeros(128, np.uint64)
    lift4[:4] = s
    assert (orc.gr_mul_gr4(k, a, s) == orc.gr_mul(k, a, lift4)).all()


# ---------------- Teichmuller lift ----------------
@pytest.mark.parametrize("k", WIDTHS)
def test_lift_fixed_point_and_one_pass(orc, rng, k):
    """The lift is a fixed point of x -> x^(2^128); one pass of 128 squarings
    already equals the reference's (ceil(log2 k)+1)-pass lift (2-adic
    valuation of (1+2z)^(2^128) - 1 is >= 129 > 64)."""
    assert (orc.lift(k, np.zeros(128, np.uint64)) == 0).all()
    assert (orc.lift(k, one()) == one()).all()
    for _ in range(2):
        x = rand_limbs(rng, k, 128)
        t = orc.lift(k, x)
        y = t.copy()
        for _ in range(128):
            y = orc.gr_mul(k, y, y)
        assert (y == t).all()
        one_pass = x.copy()
        for _ in range(128):
            one_pass = orc.gr_mul(k, one_pass, one_pass)
        assert (one_pass == t).all()
        # congruent to the natural lift mod 2 exactly when x is a unit
    for t in orc.t1(k):
        e = np.zeros(128, np.uint64)
        e[:4] = t
        assert (orc.lift(k, e) == e).all()


# ---------------- packing constants ----------------
@pytest.mark.parametrize("k", WIDTHS)
def test_lagrange_and_quotient_identities(orc, k):
    m = (1 << k) - 1
    lb = orc.consts(k, 0).reshape(16, 128)
    vb = orc.consts(k, 1)
    qb = orc.consts(k, 4).reshape(16, 16, 128)
    s = sum(lb.astype(object))
    assert [int(v) & m for v in s] == one().tolist()
    b16 = beta_pow(16, orc, k).astype(object) - beta_pow(1, orc, k).astype(object)
    assert [int(v) & m for v in b16] == vb.tolist()
    for i, j in ((0, 0), (3, 3), (2, 9), (15, 1)):
        lhs = orc.gr_mul(k, lb[i], lb[j])
        rhs = orc.gr_mul(k, qb[i, j], vb).astype(object)
        if i == j:
            rhs = rhs + lb[i].astype(object)
        assert lhs.tolist() == [int(v) & m for v in rhs]
    assert int(orc.consts(k, 5)[0]) * 15 & m == 1


@pytest.mark.parametrize("k", WIDTHS)
def test_psi_left_inverse_and_kernel(orc, rng, k):
    m = (1 << k) - 1
    a = rand_limbs(rng, k, (6, 32))
    pa = orc.pack(k, a)
    assert (orc.psi(k, pa) == a).all()
    basis = orc.consts(k, 3).reshape(32, 128)
    assert (orc.pack(k, np.eye(32, dtype=np.uint64)) == basis).all()
    assert orc.in_kernel(k, np.zeros(128, np.uint64))
    va = np.zeros(128, np.uint64)
    va[:4] = orc.consts(k, 2)
    assert orc.in_kernel(k, va) and orc.in_kernel(k, orc.consts(k, 1))
    for i in range(3):
        assert not orc.in_kernel(k, pa[i])


@pytest.mark.parametrize("k", WIDTHS)
def test_quotient_identity(orc, rng, k):
    """phi(a)phi(b) = phi(a.b) + q1*v_alpha + q2*v_beta (embedding.rs:147-167)."""
    m = (1 << k) - 1
    mm = np.uint64(m) if k < 64 else np.uint64(2**64 - 1)
    va = np.zeros(128, np.uint64)
    va[:4] = orc.consts(k, 2)
    vb = orc.consts(k, 1)
    for _ in range(3):
        a = rand_limbs(rng, k, 32)
        b = rand_limbs(rng, k, 32)
        with np.errstate(over="ignore"):
            ab = (a * b) & mm
        pa, pb, pab = orc.pack(k, np.stack([a, b, ab]))
        q1, q2 = orc.quotients(k, a, b)
        lhs = orc.gr_mul(k, pa, pb)
        with np.errstate(over="ignore"):
            rhs = (pab + orc.gr_mul(k, q1, va) + orc.gr_mul(k, q2, vb)) & mm
        assert (lhs == rhs).all()
        e = (lhs.astype(object) - pab.astype(object))
        e = np.array([int(v) & m for v in e], np.uint64)
        assert orc.in_kernel(k, e)
        ab2 = ab.copy()
        ab2[0] = (int(ab2[0]) + 1) & m
        e2 = np.array([int(v) & m for v in (lhs.astype(object) - orc.pack(k, ab2)[0].astype(object))], np.uint64)
        assert not orc.in_kernel(k, e2)

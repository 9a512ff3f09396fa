"""Pins the oracle's sumcheck layer: eq tables (partition of unity, closed
form), MLE folding (== direct multilinear evaluation), and the four sumchecks
(completeness with the lock-step sampler, soundness: a tampered first round
fails at round 0), as in piop-ring/tests/{completeness,soundness}.rs."""
import numpy as np
import pytest

from conftest import WIDTHS, rand_limbs


def gsum(xs, k):
    m = (1 << k) - 1
    return np.array([int(v) & m for v in sum(x.astype(object) for x in xs)], np.uint64)


@pytest.mark.parametrize("k", WIDTHS)
def test_eq_partition_and_closed_form(orc, rng, k):
    m = (1 << k) - 1
    r = rand_limbs(rng, k, (3, 128))
    eq = orc.eq_table(k, r)
    one = np.zeros(128, np.uint64)
    one[0] = 1
    assert (gsum(list(eq), k) == one).all()
    # eq[h1 h2 h3] = prod_i (r_i if h_i else 1 - r_i), r[0] is the MSB
    for h in (0, 5, 6):
        bits = [(h >> 2) & 1, (h >> 1) & 1, h & 1]
        acc = one
        for i, b in enumerate(bits):
            f = r[i] if b else np.array([int(v) & m for v in (one.astype(object) - r[i].astype(object))], np.uint64)
            acc = orc.gr_mul(k, acc, f)
        assert (acc == eq[h]).all()


@pytest.mark.parametrize("k", WIDTHS)
def test_fold_is_mle_evaluation(orc, rng, k):
    t = rand_limbs(rng, k, (8, 128))
    r = rand_limbs(rng, k, (3, 128))
    x = t
    for i in range(3):
        x = orc.fold(k, x, r[i])
    eq = orc.eq_table(k, r)
    direct = gsum([orc.gr_mul(k, eq[h], t[h]) for h in range(8)], k)
    assert (x[0] == direct).all()


@pytest.mark.parametrize("k", WIDTHS)
@pytest.mark.parametrize("which,ntab", [(0, 2), (1, 4), (2, 2), (3, 64)])
def test_sumcheck_complete_and_sound(orc, rng, k, which, ntab):
    nv = 3 if which != 3 else 2
    tabs = rand_limbs(rng, k, (ntab << nv, 128))
    rt = rand_limbs(rng, k, (nv, 128))
    if which == 1:  # the sub form proves sum eq*(a*b - c - e) = 0: take e = a*b - c
        N = 1 << nv
        m = (1 << k) - 1
        for h in range(N):
            ab = orc.gr_mul(k, tabs[h], tabs[N + h]).astype(object)
            tabs[3 * N + h] = np.array([int(v) & m for v in ab - tabs[2 * N + h].astype(object)], np.uint64)
    rc, _ = orc.sumcheck(k, which, nv, tabs, rt, seed=42)
    assert rc == 0
    rc, _ = orc.sumcheck(k, which, nv, tabs, rt, seed=42, tamper=True)
    assert rc == 1  # RoundCheckFailed { round: 0 }

"""GPU parity of the device-side matrix packing (setup and the SIMD prover's
per-proof grouping) on matrices that stress it: duplicated entries (one
(r, c) several times, its value split so the matrix is unchanged),
zero-valued entries, a row and a column with more than 1024 entries (the
heavy-major path of the matvecs), shuffled entry order, and an empty
matrix.  The oracle sums repeated entries as the reference's packing does
(reduce.rs:119-141, 387-397), so the proof bytes must equal the oracle's."""
import numpy as np
import pytest

import nenya_amd as na
from nenya_amd.instances import general_instance, mimd_instance, simd_instance

pytestmark = pytest.mark.gpu


def _msk(k):
    return np.uint64((1 << k) - 1) if k < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)


def _stress(m, n, k, rng, heavy=True):
    """The same matrix as m = (rows, cols, vals), written with repeats, zero
    entries, a heavy row and a heavy column, in shuffled order."""
    r, c, v = (np.asarray(x) for x in m)
    r, c, v = r.astype(np.uint32), c.astype(np.uint32), v.astype(np.uint64)
    msk = _msk(k)
    nd = min(200, len(r))
    idx = rng.choice(len(r), nd, replace=False)
    part = rng.integers(0, 1 << 62, nd, dtype=np.uint64) & msk
    v2 = v.copy()
    v2[idx] = (v[idx] - part) & msk
    rs, cs, vs = [r, r[idx]], [c, c[idx]], [v2, part]
    if heavy:
        hc = rng.integers(0, n, 1100).astype(np.uint32)
        rs += [np.full(1100, 3, np.uint32), hc]
        cs += [hc, np.full(1100, 5, np.uint32)]
        vs += [np.zeros(1100, np.uint64), np.zeros(1100, np.uint64)]
    rr, cc, vv = np.concatenate(rs), np.concatenate(cs), np.concatenate(vs)
    perm = rng.permutation(len(rr))
    return rr[perm], cc[perm], vv[perm]


@pytest.mark.parametrize("k,log_n", [(32, 11), (8, 11), (64, 11)])
def test_general_packing_with_repeats_and_heavy_majors(orc, k, log_n):
    print(f"general setup stress k={k} log_n={log_n} seed=5", flush=True)
    n = 1 << log_n
    rng = np.random.default_rng(50 + k)
    A, B, Cm, w = general_instance(log_n, k, 5)
    A2, B2, C2 = (_stress(m, n, k, rng) for m in (A, B, Cm))
    params = na.GeneralSnarkParams.secure(log_n - 5, (A2, B2, C2), 0xDEAD, 0xCAFE, k=k)
    assert set(params.setup_timing()) >= {"pcs_params", "matrices"}
    got = na.prove_general(w, params)
    want = orc.general_prove(k, log_n - 5, A, B, Cm, w)
    assert got == want
    assert na.verify_general(params, got) == "Ok"


@pytest.mark.parametrize("k", (16, 64))
def test_simd_grouping_with_repeats_and_heavy_majors(orc, k):
    print(f"simd setup stress k={k} log=11 seed=13", flush=True)
    log = 11
    n = 1 << log
    rng = np.random.default_rng(70 + k)
    A, B, Cm, wit = simd_instance(log, k, 13)
    A2, B2, C2 = (_stress(m, n, k, rng) for m in (A, B, Cm))
    params = na.SimdSnarkParams.secure(log, log, 0xDEAD, 0xCAFE, k=k)
    got = na.prove_simd(na.SimdR1csInstance(A2, B2, C2, wit), params)
    want = orc.simd_prove(k, log, log, A, B, Cm, wit)
    assert got == want
    assert na.verify_simd(na.SimdR1csInstance(A2, B2, C2, wit), params, got) == "Ok"


@pytest.mark.parametrize("k", (32, 64))
def test_mimd_packing_with_repeats_and_an_empty_instance(orc, k):
    print(f"mimd setup stress k={k} log=8 seed=17", flush=True)
    log = 8
    n = 1 << log
    rng = np.random.default_rng(90 + k)
    As, Bs, Cs, wit = mimd_instance(log, k, 17)
    As2 = [_stress(m, n, k, rng, heavy=(s == 0)) for s, m in enumerate(As)]
    Bs2 = [_stress(m, n, k, rng, heavy=False) for m in Bs]
    Cs2 = list(Cs)
    # an instance whose C matrix is written as nothing but zero entries
    # (its C is all zero in value; the witness of instance 31 stays paired
    # with the original C, so both sides prove the same zero-C instance)
    zr = np.arange(4, dtype=np.uint32)
    Cs2[31] = (zr, zr, np.zeros(4, np.uint64))
    Cs_ref = list(Cs)
    Cs_ref[31] = (np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint64))
    params = na.MimdSnarkParams.secure(log, log, (As2, Bs2, Cs2), 0xDEAD, 0xCAFE, k=k)
    got = na.prove_mimd(wit, params)
    want = orc.mimd_prove(k, log, log, As, Bs, Cs_ref, wit)
    assert got == want

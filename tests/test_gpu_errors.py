"""GPU: shape misuse that the reference rejects with assert! (simd.rs:224-231,
reduce.rs:375-377, 674-675) returns NENYA_E_SHAPE (1) through the C ABI, and
the params stay usable afterwards."""
import numpy as np
import pytest

import nenya_amd as na
from nenya_amd.instances import general_instance, mimd_instance, simd_instance

pytestmark = pytest.mark.gpu


def test_shape_errors_are_codes():
    print("shape errors", flush=True)
    A, B, Cm, w = general_instance(10, 32, 5)
    gp = na.GeneralSnarkParams.secure(5, (A, B, Cm), k=32)
    with pytest.raises(na.NenyaError) as e:
        na.prove_general(w[:-32], gp)
    assert e.value.code == 1
    assert len(na.prove_general(w, gp)) > 0

    A, B, Cm, wit = simd_instance(5, 8, 6)
    sp = na.SimdSnarkParams.secure(5, 5, k=8)
    bad = (np.append(A[0], 40).astype(np.uint32), np.append(A[1], 0).astype(np.uint32), np.append(A[2], 1))
    with pytest.raises(na.NenyaError) as e:
        na.prove_simd(na.SimdR1csInstance(bad, B, Cm, wit), sp)
    assert e.value.code == 1
    with pytest.raises(na.NenyaError) as e:
        na.prove_simd(na.SimdR1csInstance(A, B, Cm, wit[:, :16]), sp)
    assert e.value.code == 1
    assert len(na.prove_simd(na.SimdR1csInstance(A, B, Cm, wit), sp)) > 0

    As, Bs, Cs, mw = mimd_instance(5, 16, 7)
    Bs = list(Bs)
    Bs[3] = (np.append(Bs[3][0], 33).astype(np.uint32), np.append(Bs[3][1], 0).astype(np.uint32),
             np.append(Bs[3][2], 1))
    with pytest.raises(na.NenyaError) as e:
        na.MimdSnarkParams.secure(5, 5, (As, Bs, Cs), k=16)
    assert e.value.code == 1


def test_challenge_wait_time_limit_fails_the_proof(monkeypatch):
    """A device-side wait for a host-posted sumcheck challenge that runs out
    of time fails the proof with a device error instead of returning a proof
    built from a stale challenge.  The condition is forced from the host by a
    zero wait limit (every kernel still exits normally); later proofs are
    unaffected.  Exercises the single-caller pipelined round loop."""
    print("challenge wait limit", flush=True)
    A, B, Cm, w = general_instance(10, 32, 5)
    gp = na.GeneralSnarkParams.secure(5, (A, B, Cm), k=32)
    good = na.prove_general(w, gp)
    monkeypatch.setenv("NENYA_WAIT_LIMIT_US", "0")
    with pytest.raises(na.NenyaError) as e:
        na.prove_general(w, gp)
    assert e.value.code == 3
    monkeypatch.delenv("NENYA_WAIT_LIMIT_US")
    assert na.prove_general(w, gp) == good

"""The autotuner (nenya_autotune) changes only launch shapes: proofs made
after it ran are byte-identical to the ones made before, for every width and
model, and its report names a pick for every tuned size."""
import pytest

import nenya_amd as na
from nenya_amd.instances import general_instance, mimd_instance, simd_instance

pytestmark = pytest.mark.gpu


def test_proofs_after_autotune_equal_proofs_before():
    print("autotune: proofs before/after, widths 8..64", flush=True)
    made = []
    for i, (model, k, log) in enumerate([("g", 32, 14), ("g", 8, 12), ("g", 64, 11), ("s", 16, 9), ("m", 32, 7),
                                         ("s", 64, 8)]):
        if model == "g":
            A, B, Cm, w = general_instance(log, k, 90 + i)
            p = na.GeneralSnarkParams.secure(log - 5, (A, B, Cm), 0xDEAD, 0xCAFE, k=k)
            made.append((lambda p=p, w=w: na.prove_general(w, p)))
        elif model == "m":
            As, Bs, Cs, wit = mimd_instance(log, k, 90 + i)
            p = na.MimdSnarkParams.secure(log, log, (As, Bs, Cs), 0xDEAD, 0xCAFE, k=k)
            made.append((lambda p=p, wit=wit: na.prove_mimd(wit, p)))
        else:
            A, B, Cm, wit = simd_instance(log, k, 90 + i)
            p = na.SimdSnarkParams.secure(log, log, 0xDEAD, 0xCAFE, k=k)
            inst = na.SimdR1csInstance(A, B, Cm, wit)
            made.append((lambda p=p, inst=inst: na.prove_simd(inst, p)))
    before = [f() for f in made]
    rep = na.autotune(log_min=8, log_max=16)
    for k in (8, 16, 32):
        assert set(rep[f"k{k}"]["fold"]) == {f"2^{i}" for i in range(8, 17)}
        assert all(e["workgroups"] > 0 for e in rep[f"k{k}"]["gram"].values())
    assert set(rep["k64"]["fold"]) == {f"2^{i}" for i in range(8, 17)}
    after = [f() for f in made]
    assert after == before

"""Regenerates tests/golden/oracle_c0.json: digests of the oracle's outputs
on the small CPU workload (SIMD, Z/2^32, per-instance n = m = 2^5,
instance seed 7, PCS seeds 0xDEAD/0xCAFE), a General 2^10 and a MIMD Z/2^64
2^5 proof, plus sampler streams.

These fixtures pin the oracle against *itself* across changes (regression);
parity with the reference rests on the checks in tests/test_oracle_*.py.
Run:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from nenya_amd.instances import general_instance, mimd_instance, simd_instance  # noqa: E402
from oracle import load  # noqa: E402


def digest(b):
    return hashlib.blake2b(b, digest_size=16).hexdigest()


def main():
    o = load()
    out = {}
    A, B, C, wit = simd_instance(5, 32, 7)
    p = o.simd_prove(32, 5, 5, A, B, C, wit)
    out["simd_u32_log5_seed7"] = {"len": len(p), "digest": digest(p), "head": p[:64].hex()}
    A, B, C, w = general_instance(10, 32, 3)
    p = o.general_prove(32, 5, A, B, C, w)
    out["general_u32_logn10_seed3"] = {"len": len(p), "digest": digest(p), "head": p[:128].hex()}
    As, Bs, Cs, wit = mimd_instance(5, 64, 21)
    p = o.mimd_prove(64, 5, 5, As, Bs, Cs, wit)
    out["mimd_u64_log5_seed21"] = {"len": len(p), "digest": digest(p), "head": p[:128].hex()}
    out["rng_u32_seed0"] = [int(x) for x in o.rng_stream(32, 0, 0, 8)]
    out["rng_u64_seed0xdead"] = [int(x) for x in o.rng_stream(64, 0xDEAD, 1, 8)]
    out["raa_perm1_k16_seed0xdead"] = [int(x) for x in o.raa_setup(32, 16, 4, 0xDEAD)[0]]
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_c0.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()

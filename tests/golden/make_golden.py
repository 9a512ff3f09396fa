"""Regenerates the golden fixtures from the oracle (the C restatement of the
reference prover, oracle/).

tests/golden/oracle_c0.json: digests of the oracle's outputs on the small
CPU workload (SIMD, Z/2^32, per-instance n = m = 2^5, instance seed 7, PCS
seeds 0xDEAD/0xCAFE), a General 2^10 and a MIMD Z/2^64 2^5 proof, plus
sampler streams.

tests/golden/oracle_bench.json (--bench): digests of the oracle's proofs at
the sizes the benchmark times -- General Z/2^32 at 2^20 constraints (instance
seed 1), SIMD at n = m = 2^17 per instance for k = 8, 16, 32, 64 (instance
seed 301 + k) and the PNG-filter batch of 2^20 bytes (gradient images) --
so the GPU proofs at those sizes are checked byte for byte without running
the oracle on the GPU host.  The oracle runs on every host core (OpenMP);
expect a few minutes.

These fixtures pin the GPU path to the oracle at full size, and the oracle
against itself across changes; parity of the oracle with the reference rests
on the checks in tests/test_oracle_*.py.
Run:  python tests/golden/make_golden.py [--bench]
"""
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from nenya_amd.instances import general_instance, mimd_instance, png_instance, simd_instance  # noqa: E402
from oracle import load  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def digest(b):
    return hashlib.blake2b(b, digest_size=16).hexdigest()


def small():
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
    with open(os.path.join(HERE, "oracle_c0.json"), "w") as f:
        json.dump(out, f, indent=1)


def bench():
    o = load()
    out = {"_about": "oracle proof digests (blake2b-128 of the serialized proof) at the benchmark sizes; "
                     "PCS seeds 0xDEAD / 0xCAFE"}

    def rec(name, p, t0, **kw):
        out[name] = dict(kw, len=len(p), digest=digest(p), head=p[:64].hex())
        print(f"{name}: {len(p)} bytes in {time.perf_counter() - t0:.1f} s", flush=True)

    t0 = time.perf_counter()
    A, B, C, w = general_instance(20, 32, 1)
    rec("general_u32_logn20_seed1", o.general_prove(32, 15, A, B, C, w), t0, model="general", k=32, log_n=20, seed=1)
    del A, B, C, w
    for k in (8, 16, 32, 64):
        t0 = time.perf_counter()
        A, B, C, wit = simd_instance(17, k, 301 + k)
        rec(f"simd_u{k}_log17_seed{301 + k}", o.simd_prove(k, 17, 17, A, B, C, wit), t0, model="simd", k=k, log=17,
            seed=301 + k)
        del A, B, C, wit
    t0 = time.perf_counter()
    ln, lm, A, B, C, wit = png_instance(20, "gradient")
    rec("png_u8_logbytes20_gradient", o.simd_prove(8, ln, lm, A, B, C, wit), t0, model="simd", k=8, log_n=ln,
        log_m=lm, image="gradient")
    with open(os.path.join(HERE, "oracle_bench.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    if "--bench" in sys.argv:
        bench()
    else:
        small()

"""MI355X-native Nenya prover (gfx950 HIP kernels behind a C ABI).

Python mirror of the reference prover interface (tnyuzg/nenya,
crates/nenya/snark-ring/src/lib.rs:23-36):

    params = SimdSnarkParams.secure(log_n, log_m, seed_w, seed_q, k=32)
    proof  = prove_simd(SimdR1csInstance(a, b, c, witnesses), params)

    params = MimdSnarkParams.secure(log_n, log_m, (As, Bs, Cs), seed_w, seed_e, k=64)
    proof  = prove_mimd(witnesses, params)       # 32 independent instances

    params = GeneralSnarkParams.secure(log_blocks, (a, b, c), seed_w, seed_e, k=32)
    proof  = prove_general(w, params)

Batched / served proofs (several proofs overlapping on one GPU):

    srv = Server(workers=4)
    job = srv.submit_general(params, w)       # or submit_simd / submit_mimd
    proof, queued_ms, run_ms = srv.wait(job)

Verifiers (host checks, row recompute on the GPU) return "Ok" or the
reference's VerifyError name:

    verify_general(params, proof)   # verify_simd(inst, params, proof), verify_mimd(params, proof)

Matrices are COO triples (rows u32, cols u32, vals) with vals limbs of Z/2^k;
witnesses are limb arrays.  Proofs are the reference's serialized bytes.
Every call runs the HIP library nenya_amd/libnenya_amd.so; there is no CPU
fallback: without the library or a GPU the calls raise.
"""
from ._lib import (  # noqa: F401
    DeviceArray, NenyaError, autotune, GeneralSnarkParams, MimdSnarkParams, Server, SimdR1csInstance, SimdSnarkParams,
    device_copy, device_count, device_synchronize, lib_path, proof_container, proof_payload, prove_general, prove_general_device, prove_mimd, prove_simd,
    verify_general, verify_mimd, verify_simd,
)

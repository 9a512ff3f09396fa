"""Profiling driver: proves one named workload a few times, for runs under
rocprofv3 (kernel traces and hardware counters of a known workload).

  python -m nenya_amd.profile_run simd K LOG [--proofs N] [--seed S]
      SIMD proof, n = m = 2^LOG per instance, Z/2^K (the width sweep's
      instance: seed 301 + K by default)
  python -m nenya_amd.profile_run general K LOG_N [--proofs N] [--seed S]
      General proof at 2^LOG_N constraints (instance seed 1 by default)
  --copy-calib  first runs a 256 MiB device-to-device copy through the
      library's copy kernel (the known-size transfer for FETCH_SIZE /
      WRITE_SIZE calibration)

Each proof's bytes are checked with the library's verifier; one line per
proof goes to stdout.
"""
import argparse
import sys
import time

import numpy as np

import nenya_amd as na
from nenya_amd.instances import general_instance, simd_instance

_DT = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("model", choices=("simd", "general"))
    ap.add_argument("k", type=int)
    ap.add_argument("log", type=int)
    ap.add_argument("--proofs", type=int, default=3)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--copy-calib", action="store_true")
    a = ap.parse_args(argv)
    dt = _DT[a.k]
    if a.copy_calib:
        src = na.DeviceArray(np.zeros(256 << 20, np.uint8))
        dst = na.DeviceArray(np.zeros(256 << 20, np.uint8))
        na.device_copy(dst, src)
        print("copy calibration: 268435456 bytes device-to-device", flush=True)
        del src, dst
    if a.model == "simd":
        seed = 301 + a.k if a.seed is None else a.seed
        A, B, C, wit = simd_instance(a.log, a.k, seed)
        p = na.SimdSnarkParams.secure(a.log, a.log, 0xDEAD, 0xCAFE, k=a.k)
        inst = na.SimdR1csInstance((A[0], A[1], A[2].astype(dt)), (B[0], B[1], B[2].astype(dt)),
                                   (C[0], C[1], C[2].astype(dt)), np.ascontiguousarray(wit.astype(dt)))

        def prove():
            return na.prove_simd(inst, p, view=True)

        def verify(b):
            return na.verify_simd(inst, p, b)
    else:
        seed = 1 if a.seed is None else a.seed
        A, B, C, w = general_instance(a.log, a.k, seed)
        p = na.GeneralSnarkParams.secure(a.log - 5, (A, B, C), 0xDEAD, 0xCAFE, k=a.k)
        wd = na.DeviceArray(w.astype(dt))

        def prove():
            return na.prove_general_device(wd.ptr, wd.size, p, view=True)

        def verify(b):
            return na.verify_general(p, b)
    for i in range(a.proofs):
        t = time.perf_counter()
        b = bytes(prove())
        ms = (time.perf_counter() - t) * 1e3
        print(f"{a.model} u{a.k} 2^{a.log} proof {i}: {len(b)} bytes, {ms:.2f} ms (with copy-out), "
              f"verifier {verify(b)}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

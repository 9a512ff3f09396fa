#!/usr/bin/env python3
"""Headline benchmark: prover seconds for a General-model proof over Z/2^32 at
2^20 R1CS constraints (log_blocks = 15) on one MI355X.

A step is one full proof (prove_general on a witness already resident in
HBM) whose serialized bytes land in a pinned host buffer.  Instance: the
random half-constrained R1CS of nenya_amd.instances (seed 1), PCS seeds
0xDEAD / 0xCAFE, delta = 0.19, lambda = 100, rate 1/4, t = 1060.

Roofline: one profiled proof (HIP events around every launch) names the
dominant throughput kernel -- the one with the largest summed time among
kernels moving at least 1 MiB per launch; the timed proofs then carry events
around that kernel's launches only, and its algorithmic bytes (or int8 MFMA
operations) per launch over its average duration give `achieved`.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
INT8_PEAK_TOPS = 5000.0    # MI355X dense int8 MFMA, TOP/s
MX_KERNELS = ("k_fold_mx", "k_gram_lin", "k_mulfix_mx", "k_cgemm", "k_scale_mx")  # int8 digit-plane GEMMs


def log(*a):
    print(*a, file=sys.stderr, flush=True)


GOLDEN = os.path.join(ROOT, "tests", "golden", "oracle_bench.json")


def golden(name):
    """The oracle's digest record for a benchmark workload, if committed."""
    try:
        return json.load(open(GOLDEN)).get(name)
    except (OSError, ValueError):
        return None


def digest(b):
    import hashlib
    return hashlib.blake2b(bytes(b), digest_size=16).hexdigest()


def check_proof(proof, gold, verdict):
    """Verification record of a timed proof: the library verifier's verdict
    and, where the oracle's digest for this workload is committed, whether
    the bytes match it."""
    out = {"verifier": verdict}
    if gold is not None:
        out["oracle_digest_match"] = len(proof) == gold["len"] and digest(proof) == gold["digest"]
    out["ok"] = verdict == "Ok" and out.get("oracle_digest_match", True)
    return out


def cpu_baseline(k, log_n, seed):
    """The oracle (the C restatement of the reference prover's algorithm,
    including the per-triple quotients and the reference's lift) proving the
    same workload on the host's cores (OpenMP threads over its loops), timed
    once; its bytes are checked against the committed digest."""
    from oracle import load
    from nenya_amd.instances import general_instance
    o = load()
    A, B, C, w = general_instance(log_n, k, seed)
    t = time.perf_counter()
    p = o.general_prove(k, log_n - 5, A, B, C, w)
    dt = time.perf_counter() - t
    g = golden(f"general_u{k}_logn{log_n}_seed{seed}")
    return {"value": round(dt, 3), "unit": "s", "cores": o.threads(), "nproc": os.cpu_count(), "kind": "port",
            "sample": f"one oracle prove_general at 2^{log_n} u{k} constraints (the full workload, not scaled), "
                      f"{o.threads()} OpenMP threads",
            "oracle_digest_match": (g is not None and digest(p) == g["digest"]) if g else None}


_LIMB = {8: "uint8", 16: "uint16", 32: "uint32", 64: "uint64"}


def as_limbs(x, k):
    """Input arrays in the proof's limb type (set-up work, outside the timed
    region: the instance generators draw uint64 values)."""
    import numpy as np
    return np.ascontiguousarray(np.asarray(x).astype(_LIMB[k]))


def coo_limbs(m, k):
    r, c, v = m
    return (r, c, as_limbs(v, k))


def mixed_jobs(na, njobs, seed):
    """The mixed batch: logN ~ U{12..20}, k ~ U{8,16,32,64}, model ~ U{SIMD,
    MIMD, General} (SIMD/MIMD per-instance log = logN - 5, General
    log_blocks = logN - 5), instance seed = seed + i."""
    import numpy as np
    from nenya_amd.instances import general_instance, mimd_instance, simd_instance
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(njobs):
        log_n, k, model = int(rng.integers(12, 21)), int(rng.choice([8, 16, 32, 64])), int(rng.integers(0, 3))
        if model == 1:
            A, B, C, w = general_instance(log_n, k, seed + i)
            p = na.GeneralSnarkParams.secure(log_n - 5, (A, B, C), 0xDEAD, 0xCAFE, k=k)
            jobs.append(("general", p, as_limbs(w, k)))
        elif model == 2:
            As, Bs, Cs, wit = mimd_instance(log_n - 5, k, seed + i)
            p = na.MimdSnarkParams.secure(log_n - 5, log_n - 5, (As, Bs, Cs), 0xDEAD, 0xCAFE, k=k)
            jobs.append(("mimd", p, as_limbs(wit, k)))
        else:
            A, B, C, wit = simd_instance(log_n - 5, k, seed + i)
            p = na.SimdSnarkParams.secure(log_n - 5, log_n - 5, 0xDEAD, 0xCAFE, k=k)
            jobs.append(("simd", p, na.SimdR1csInstance(coo_limbs(A, k), coo_limbs(B, k), coo_limbs(C, k),
                                                        as_limbs(wit, k))))
    return jobs


def mimd_serving_jobs(na, njobs, seed):
    """The MIMD serving mix: k = 64, logN ~ U{10..22} (per-instance log =
    logN - 5, A rows with 1..4 nonzeros); one circuit (32 instance matrix
    sets) per size, set up before the run, and a fresh satisfying witness
    batch per job (seed + i)."""
    import numpy as np
    from nenya_amd.instances import mimd_instance, mimd_witnesses
    rng = np.random.default_rng(seed)
    sizes = [int(rng.integers(10, 23)) for _ in range(njobs)]
    circ = {}
    jobs = []
    for i, log_n in enumerate(sizes):
        lg = log_n - 5
        if lg not in circ:
            As, Bs, Cs, _ = mimd_instance(lg, 64, seed + 1000 * lg, max_a=4)
            circ[lg] = (na.MimdSnarkParams.secure(lg, lg, (As, Bs, Cs), 0xDEAD, 0xCAFE, k=64), As, Bs)
        p, As, Bs = circ[lg]
        jobs.append(("mimd", p, as_limbs(mimd_witnesses(As, Bs, lg, 64, seed + i), 64)))
    return jobs, sorted(set(sizes))


def time_proofs(fn, reps):
    """Median ms per proof over `reps` proofs after one warm-up (garbage of
    the previous section collected first, so a host pause cannot skew it)."""
    import gc
    gc.collect()
    fn()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    ts.sort()
    return ts[len(ts) // 2] * 1e3


def width_sweep(na, seed, reps=5):
    """SIMD at n = m = 2^17 per instance for k = 8, 16, 32, 64, and the PNG
    filter batch (u8, 32 images x 2^15 bytes) -- ms per proof, host inputs,
    the proof left in the params' pinned buffer (a view, no copy)."""
    from nenya_amd.instances import png_instance, simd_instance
    out, ver = {}, {}
    for k in (8, 16, 32, 64):
        A, B, C, wit = simd_instance(17, k, seed + k)
        p = na.SimdSnarkParams.secure(17, 17, 0xDEAD, 0xCAFE, k=k)
        inst = na.SimdR1csInstance(coo_limbs(A, k), coo_limbs(B, k), coo_limbs(C, k), as_limbs(wit, k))
        out[f"simd_u{k}_2p17_ms"] = round(time_proofs(lambda: na.prove_simd(inst, p, view=True), reps), 2)
        last = bytes(na.prove_simd(inst, p, view=True))
        ver[f"simd_u{k}"] = check_proof(last, golden(f"simd_u{k}_log17_seed{seed + k}"), na.verify_simd(inst, p, last))
        log(f"  simd u{k} 2^17: {out[f'simd_u{k}_2p17_ms']} ms/proof  {ver[f'simd_u{k}']}")
        del p, inst
    ln, lm, A, B, C, wit = png_instance(20, "gradient")
    p = na.SimdSnarkParams.secure(ln, lm, 0xDEAD, 0xCAFE, k=8)
    inst = na.SimdR1csInstance(coo_limbs(A, 8), coo_limbs(B, 8), coo_limbs(C, 8), as_limbs(wit, 8))
    out["png_u8_2p20_bytes_ms"] = round(time_proofs(lambda: na.prove_simd(inst, p, view=True), reps), 2)
    last = bytes(na.prove_simd(inst, p, view=True))
    ver["png_u8"] = check_proof(last, golden("png_u8_logbytes20_gradient"), na.verify_simd(inst, p, last))
    out["proofs"] = "left in the params' pinned buffer (view), as the headline's"
    out["png_shape"] = {"n": 2 ** ln, "m": 2 ** lm,
                        "images": "32 per-row linear gradients + uniform +-16 noise (nenya_amd.instances.png_instance)"}
    out["verified"] = ver
    log(f"  png u8 log_bytes=20 (n=2^{ln}, m=2^{lm}): {out['png_u8_2p20_bytes_ms']} ms/proof  {ver['png_u8']}")
    return out


def run_queue(srv, jobs, arrivals=None):
    """Submit `jobs` to the server `srv` (all at t=0, or at the given arrival
    offsets in s); returns (wall s, latencies ms).  Proofs stay in their
    pinned host buffers (no copy into Python objects)."""
    import numpy as np
    t0 = time.perf_counter()
    ids, sub = [], []
    for i, (model, p, x) in enumerate(jobs):
        if arrivals is not None:
            while time.perf_counter() - t0 < arrivals[i]:
                time.sleep(0.0002)
        sub.append(time.perf_counter())
        if model == "general":
            ids.append(srv.submit_general(p, x))
        elif model == "mimd":
            ids.append(srv.submit_mimd(p, x))
        else:
            ids.append(srv.submit_simd(x, p))
    lat = []
    for jid in ids:
        _, q, r = srv.wait(jid, copy=False)
        lat.append(q + r)
    wall = time.perf_counter() - t0
    return wall, np.array(lat)


def prove_direct(na, job):
    """One job through the drop-in entry points (nenya_prove_*_view: the
    single-proof path with pipelined rounds and forked commitments), the
    proof left in its params' pinned buffer."""
    model, p, x = job
    if model == "general":
        return na.prove_general(x, p, view=True)
    if model == "mimd":
        return na.prove_mimd(x, p, view=True)
    return na.prove_simd(x, p, view=True)


def run_plain(na, jobs, arrivals=None):
    """The plain path: the jobs proved one at a time, in order, by direct
    calls (all available at t=0, or each from its arrival offset in s on);
    returns (wall s, latencies ms from each job's arrival to its proof)."""
    import numpy as np
    t0 = time.perf_counter()
    lat = []
    for i, job in enumerate(jobs):
        a = 0.0 if arrivals is None else arrivals[i]
        while time.perf_counter() - t0 < a:
            time.sleep(0.0002)
        prove_direct(na, job)
        lat.append((time.perf_counter() - t0 - a) * 1e3)
    return time.perf_counter() - t0, np.array(lat)


def verify_queue(na, jobs, workers):
    """One more (untimed) pass of the queue keeping the proofs, each checked
    with the library's verifier; returns the number accepted."""
    srv = na.Server(workers=workers)
    ids = []
    for model, p, x in jobs:
        if model == "general":
            ids.append(srv.submit_general(p, x))
        elif model == "mimd":
            ids.append(srv.submit_mimd(p, x))
        else:
            ids.append(srv.submit_simd(x, p))
    ok = 0
    for (model, p, x), jid in zip(jobs, ids):
        proof, _, _ = srv.wait(jid)
        if model == "general":
            v = na.verify_general(p, proof)
        elif model == "mimd":
            v = na.verify_mimd(p, proof)
        else:
            v = na.verify_simd(x, p, proof)
        ok += v == "Ok"
    srv.close()
    return ok


def queue_stats(na, jobs, workers, seed):
    """Batch (all jobs at t=0) and serving (Poisson arrivals at 70% of the
    batch throughput) figures for the server with `workers` streams and for
    the plain path: the same jobs proved one at a time by direct calls to
    the drop-in entry points.  Each path first proves the whole job list
    once, untimed, so device pools and pinned proof buffers have their
    steady-state sizes, as in a long-running process."""
    import numpy as np

    def stats(wall, lat, batch):
        d = {"p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2)}
        if batch:
            d = {"proofs_per_s": round(len(jobs) / wall, 2), **d}
        return d

    out, serve = {}, {}
    srv = na.Server(workers=workers)
    try:
        run_queue(srv, jobs)
        out["batched"] = stats(*run_queue(srv, jobs), True)
        log(f"  batch batched  workers={workers}: {out['batched']}")
        lam = 0.7 * out["batched"]["proofs_per_s"]
        arr = np.cumsum(np.random.default_rng(seed).exponential(1.0 / lam, len(jobs)))
        serve["lambda_per_s"] = round(lam, 2)
        serve["batched"] = stats(*run_queue(srv, jobs, arrivals=arr), False)
        log(f"  serve batched  workers={workers} lambda={lam:.1f}/s: {serve['batched']}")
    finally:
        srv.close()
    run_plain(na, jobs)
    out["plain"] = stats(*run_plain(na, jobs), True)
    log(f"  batch plain (direct calls): {out['plain']}")
    serve["plain"] = stats(*run_plain(na, jobs, arrivals=arr), False)
    log(f"  serve plain (direct calls) lambda={lam:.1f}/s: {serve['plain']}")
    out["ratio"] = round(out["batched"]["proofs_per_s"] / out["plain"]["proofs_per_s"], 3)
    return out, serve


def traffic_for(kernel, k, log_n):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE scaled by a known-size copy), if one
    exists for this kernel and workload."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        d = json.load(open(p))
    except (OSError, ValueError):
        return None
    e = d.get(f"{kernel}@u{k}_2p{log_n}")
    return e.get("bytes_per_launch") if e else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log-n", type=int, default=20)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-log-n", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-jobs", type=int, default=64, help="mixed batch size (0 skips batch/serve figures)")
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the width sweep, the MIMD serving mix and the 2^25 General proof")
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in launch shapes")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a 256 MiB device copy (PMC calibration of FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); independent proofs shard "
              "trivially across GPUs (one process per GPU, jobs placed by device footprint)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import nenya_amd as na
    from nenya_amd.instances import general_instance

    k, log_n = args.k, args.log_n
    t0 = time.perf_counter()
    A, B, C, w = general_instance(log_n, k, args.seed)
    log(f"instance 2^{log_n} u{k}: nnz A={len(A[0])} B={len(B[0])} C={len(C[0])} ({time.perf_counter() - t0:.1f}s)")
    t0 = time.perf_counter()
    params = na.GeneralSnarkParams.secure(log_n - 5, (A, B, C), 0xDEAD, 0xCAFE, k=k)
    setup = {"total_ms": round((time.perf_counter() - t0) * 1e3, 1),
             **{k2: round(v, 1) for k2, v in params.setup_timing().items()}}
    log(f"setup (matrix packing + PCS params): {setup}")
    wd = na.DeviceArray(as_limbs(w, k))
    na.device_synchronize()
    # launch shapes of the sumcheck kernels picked by the autotuner (every
    # width, fold / Gram work sizes 2^8 .. 2^20); all later proofs use them
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        rep = na.autotune(log_min=8, log_max=20)
        tune = {"seconds": round(time.perf_counter() - t0, 3), "picks": {}}
        for kk, kernels in rep.items():
            if kk == "seconds":
                continue
            tune["picks"][kk] = {kern: {sz: (f"{e['pick']}/{e['workgroups']}" if "pick" in e else e["workgroups"])
                                        for sz, e in d.items()} for kern, d in kernels.items()}
            tune["picks"][kk]["us_saved_vs_builtin"] = {
                kern: round(sum(max(0.0, e["builtin_us"] - e["us"]) for e in d.values()), 1) for kern, d in kernels.items()}
        log(f"autotune: {tune['seconds']} s")
    n = len(w)

    if args.copy_calib:
        src = na.DeviceArray(np.zeros(256 << 20, np.uint8))
        dst = na.DeviceArray(np.zeros(256 << 20, np.uint8))
        na._lib.device_copy(dst, src)
        na.device_synchronize()
        del src, dst
        log("copy calibration: 268435456 bytes device-to-device")

    proofs = []
    for i in range(args.warmup):
        proofs.append(na.prove_general_device(wd.ptr, n, params))
        log(f"warmup {i}: proof {len(proofs[-1])} bytes")

    # profiled proof: per-kernel HIP-event durations; then a proof with the
    # stage breakdown alone (events at stage boundaries only, so the stages
    # show the proof's own overlapped schedule)
    params.set_profile(True)
    na.prove_general_device(wd.ptr, n, params)
    stats = params.kernel_stats()
    params.set_profile(False)
    params.set_timing(True)
    na.prove_general_device(wd.ptr, n, params)
    stages = params.last_timing()
    params.set_timing(False)
    for name, s in sorted(stats.items(), key=lambda kv: -kv[1]["ms"]):
        log(f"  {name:16s} {s['ms']:9.3f} ms  x{s['count']:<5d} {s['bytes'] / max(s['ms'], 1e-9) / 1e6:9.1f} GB/s")
    for name, v in stages.items():
        log(f"  stage {name:16s} {v:9.3f} ms")
    top_name, top = max(stats.items(), key=lambda kv: kv[1]["ms"])
    thr = {nm: s for nm, s in stats.items() if s["bytes"] / s["count"] >= (1 << 20)}
    dom_name = max(thr.items(), key=lambda kv: kv[1]["ms"])[0]

    # timed proofs; HIP events bracket the dominant kernel's launches in the
    # last timed proof only (each event record adds ~6 us of queue gap, so
    # events on every launch of every proof would slow the proofs measured)
    params.set_profile(False)
    params.set_profile_only(dom_name)
    na.device_synchronize()
    t = time.perf_counter()
    for i in range(args.steps):
        if i == args.steps - 1:
            params.set_profile(True)
        # the proof lands in a pinned host buffer owned by the params (no copy)
        pv = na.prove_general_device(wd.ptr, n, params, view=True)
    na.device_synchronize()
    elapsed = time.perf_counter() - t
    dom = params.kernel_stats()[dom_name]
    params.set_profile(False)
    params.set_profile_only(None)
    ms = elapsed / args.steps * 1e3
    p = bytes(pv)
    log(f"timed {args.steps} proofs: {ms:.2f} ms/proof")
    if proofs and p != proofs[0]:
        raise SystemExit("non-deterministic proof bytes")
    verified = check_proof(p, golden(f"general_u{k}_logn{log_n}_seed{args.seed}"), na.verify_general(params, p))
    log(f"headline proof: {verified}")

    # the drop-in boundary: host witness in, owned proof bytes out
    # (nenya_prove_general, as a Rust caller of prove_general would use it),
    # and the same with the proof left as a view of the pinned buffer
    wh = as_limbs(w, k)
    dropin_ms = time_proofs(lambda: na.prove_general(wh, params, copy=False).close(), 5)
    dropin_bytes_ms = time_proofs(lambda: na.prove_general(wh, params), 5)
    dropin_view_ms = time_proofs(lambda: na.prove_general(wh, params, view=True), 5)
    log(f"drop-in (host witness, owned proof): {dropin_ms:.2f} ms/proof; as Python bytes {dropin_bytes_ms:.2f} ms; "
        f"host witness, view: {dropin_view_ms:.2f} ms")

    per_launch_ms = dom["ms"] / dom["count"]
    per_launch_bytes = dom["bytes"] / dom["count"]
    gbs = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
    nd = k // 8
    int8_ops = dom["macs"] / dom["count"] * (nd * (nd + 1) // 2) * 2 if dom_name in MX_KERNELS else 0.0
    tops = int8_ops / (per_launch_ms * 1e-3) / 1e12
    traffic = traffic_for(dom_name, k, log_n)
    if int8_ops and tops / INT8_PEAK_TOPS > gbs / HBM_PEAK_GBS:
        roofline = {"bound": "mfma", "achieved": round(tops, 2), "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s",
                    "frac": round(tops / INT8_PEAK_TOPS, 4), "traffic": traffic}
    else:
        roofline = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic}
    roofline.update({"kernel": dom_name, "avg_launch_ms": round(per_launch_ms, 4),
                     "launches_per_proof": dom["count"], "event_timed_proofs": 1,
                     "algorithmic_bytes_per_launch": round(per_launch_bytes),
                     "int8_mfma_tops": round(tops, 2) if int8_ops else None,
                     "top_kernel_by_time": top_name, "top_kernel_us_per_launch": round(top["ms"] / top["count"] * 1e3, 1)})

    batch = serve = None
    if args.batch_jobs > 0:
        t0 = time.perf_counter()
        jobs = mixed_jobs(na, args.batch_jobs, 1000 + args.seed)
        log(f"mixed batch: {len(jobs)} jobs set up in {time.perf_counter() - t0:.1f}s")
        q, srv = queue_stats(na, jobs, args.workers, 77 + args.seed)
        nver = verify_queue(na, jobs, args.workers)
        log(f"  batch proofs accepted by the verifier: {nver}/{len(jobs)}")
        batch = {"jobs": len(jobs), "workers": args.workers, "proofs_per_s": q["batched"]["proofs_per_s"],
                 "verified": f"{nver}/{len(jobs)}",
                 "p50_ms": q["batched"]["p50_ms"], "p99_ms": q["batched"]["p99_ms"], "plain": q["plain"],
                 "ratio_vs_plain": q["ratio"],
                 "plain_path": "the same jobs one at a time through nenya_prove_*_view (direct calls)",
                 "mix": "logN~U{12..20}, k~U{8,16,32,64}, model~U{SIMD,MIMD,General}, host inputs, seed_i=base+i",
                 "warm": "each path (server and direct calls) proves the job list once, untimed, before its timed pass"}
        serve = {"lambda_per_s": srv["lambda_per_s"], "jobs": len(jobs), "workers": args.workers,
                 "p50_ms": srv["batched"]["p50_ms"], "p99_ms": srv["batched"]["p99_ms"], "plain": srv["plain"],
                 "arrivals": "Poisson at 70% of the batched throughput, same job mix"}
        del jobs

    widths = large = None
    if not args.no_extras:
        log("width sweep:")
        widths = width_sweep(na, 300 + args.seed)
        t0 = time.perf_counter()
        mj, sizes = mimd_serving_jobs(na, 32, 2000 + args.seed)
        log(f"mimd serving mix: {len(mj)} jobs over per-instance logs {sizes} set up in {time.perf_counter() - t0:.1f}s")
        # setup stages of the largest circuit of the mix
        pbig = max((j[1] for j in mj), key=lambda pp: pp.log_n)
        mimd_setup = {"log_n": pbig.log_n, **{k2: round(v, 1) for k2, v in pbig.setup_timing().items()}}
        log(f"  mimd u64 setup of the largest circuit: {mimd_setup}")
        q, srv = queue_stats(na, mj, args.workers, 99 + args.seed)
        nver = verify_queue(na, mj, args.workers)
        log(f"  mimd serving proofs accepted by the verifier: {nver}/{len(mj)}")
        large = {"mimd_u64_serve": {"jobs": len(mj), "workers": args.workers, "lambda_per_s": srv["lambda_per_s"],
                                    "verified": f"{nver}/{len(mj)}",
                                    "p50_ms": srv["batched"]["p50_ms"], "p99_ms": srv["batched"]["p99_ms"],
                                    "plain": srv["plain"], "batch": q, "setup_ms": mimd_setup,
                                    "mix": "k=64, logN~U{10..22}, one circuit per size, witness per job"}}
        del mj
        del params, wd
        t0 = time.perf_counter()
        A2, B2, C2, w2 = general_instance(25, 32, args.seed)
        p2 = na.GeneralSnarkParams.secure(20, (A2, B2, C2), 0xDEAD, 0xCAFE, k=32)
        setup25 = {"total_ms": round((time.perf_counter() - t0) * 1e3, 1),
                   **{k2: round(v, 1) for k2, v in p2.setup_timing().items()}}
        large["general_u32_2p25_setup_ms"] = setup25
        w2 = as_limbs(w2, 32)
        log(f"general u32 2^25: set up {setup25} (instance generation included in total_ms)")
        # three proofs, each timed alone, copied out and verified after its timing
        ts, checks = [], []
        for _ in range(3):
            t = time.perf_counter()
            pv25 = na.prove_general(w2, p2, view=True)
            ts.append((time.perf_counter() - t) * 1e3)
            pf25 = bytes(pv25)
            del pv25
            checks.append(dict(check_proof(pf25, None, na.verify_general(p2, pf25)), digest=digest(pf25)))
            del pf25
        ms25 = sorted(ts[1:])[0] if len(ts) > 1 else ts[0]
        large["general_u32_2p25_ms"] = round(ms25, 1)
        large["general_u32_2p25_ms_each"] = [round(x, 1) for x in ts]
        large["general_u32_2p25_verified"] = checks
        log(f"  general u32 2^25: {[round(x, 1) for x in ts]} ms/proof (host witness) {checks}")
        del p2, A2, B2, C2, w2

    cpu = None
    if not args.no_cpu_baseline:
        log(f"cpu baseline: oracle at 2^{args.cpu_log_n} ...")
        cpu = cpu_baseline(k, args.cpu_log_n, args.seed)
        log(f"cpu baseline {cpu['value']:.1f} s on {cpu['cores']} threads")

    out = {
        "metric": "prover sec @2^20 u32 constraints; batched proofs/s + p99 vs plain GPU path",
        "value": round(ms / 1e3, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"u{k}",
        "data": "synthetic",
        "config": {"workload": f"General-model proof, Z/2^{k}, 2^{log_n} constraints (log_blocks {log_n - 5}), "
                               f"random R1CS seed {args.seed}, witness resident in HBM, proof left as a view of "
                               "the params' pinned host buffer",
                   "constraints": 2 ** log_n, "log_blocks": log_n - 5, "k": k, "proof_bytes": len(p),
                   "proofs_per_s": round(1e3 / ms, 3), "verified": verified,
                   "dropin_ms": round(dropin_ms, 3), "dropin_bytes_ms": round(dropin_bytes_ms, 3),
                   "dropin_view_ms": round(dropin_view_ms, 3),
                   "dropin": "nenya_prove_general: host witness uploaded, owned proof returned and released with "
                             "nenya_free; dropin_bytes_ms also copies it into a Python bytes object; "
                             "dropin_view_ms leaves it as a view of the params' buffer"},
        "roofline": roofline,
        "autotune": tune,
        "cpu_baseline": cpu,
        "stages_ms": {k2: round(v, 3) for k2, v in stages.items()},
        "setup_ms": setup,
        "batch": batch,
        "serve": serve,
        "widths": widths,
        "large": large,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

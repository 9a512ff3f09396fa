"""GPU: proofs made through the batch/serve entry points (several worker
streams, mixed widths and SIMD / MIMD / General models in one queue) are byte-identical to the
direct path's, and the oracle verifier accepts them."""
import numpy as np
import pytest

import nenya_amd as na
from nenya_amd.instances import general_instance, mimd_instance, simd_instance

pytestmark = pytest.mark.gpu


def test_mixed_queue_matches_direct_path(orc):
    print("server: 3 workers, 11 mixed jobs", flush=True)
    jobs = []
    specs = [("g", 32, 10), ("s", 8, 6), ("m", 64, 6), ("g", 16, 10), ("s", 64, 5), ("g", 8, 11), ("m", 8, 7),
             ("s", 32, 7), ("g", 64, 10), ("m", 32, 5), ("g", 32, 11)]
    for i, (model, k, log) in enumerate(specs):
        if model == "g":
            A, B, Cm, w = general_instance(log, k, 40 + i)
            p = na.GeneralSnarkParams.secure(log - 5, (A, B, Cm), 0xDEAD, 0xCAFE, k=k)
            jobs.append(("g", k, log, p, (A, B, Cm), w))
        elif model == "m":
            As, Bs, Cs, wit = mimd_instance(log, k, 40 + i)
            p = na.MimdSnarkParams.secure(log, log, (As, Bs, Cs), 0xDEAD, 0xCAFE, k=k)
            jobs.append(("m", k, log, p, (As, Bs, Cs), wit))
        else:
            A, B, Cm, wit = simd_instance(log, k, 40 + i)
            p = na.SimdSnarkParams.secure(log, log, 0xDEAD, 0xCAFE, k=k)
            jobs.append(("s", k, log, p, (A, B, Cm), wit))
    srv = na.Server(workers=3)
    ids = []
    for model, k, log, p, m3, w in jobs:
        if model == "g":
            ids.append(srv.submit_general(p, w))
        elif model == "m":
            ids.append(srv.submit_mimd(p, w))
        else:
            ids.append(srv.submit_simd(na.SimdR1csInstance(*m3, w), p))
    for jid, (model, k, log, p, m3, w) in zip(ids, jobs):
        got, queued, run = srv.wait(jid)
        assert run > 0 and queued >= 0
        if model == "g":
            direct = na.prove_general(w, p)
            assert got == direct, f"general k={k} log={log}"
            assert orc.general_verify(k, log - 5, *m3, got) == "Ok"
        elif model == "m":
            direct = na.prove_mimd(w, p)
            assert got == direct, f"mimd k={k} log={log}"
            assert orc.mimd_verify(k, log, log, *m3, got) == "Ok"
        else:
            direct = na.prove_simd(na.SimdR1csInstance(*m3, w), p)
            assert got == direct, f"simd k={k} log={log}"
            assert orc.simd_verify(k, log, log, *m3, got) == "Ok"
    srv.close()


def test_concurrent_small_folds_match_direct_path():
    """Twelve proofs over two shared params objects on four workers.  The
    last sumcheck rounds fold a few tiles in place with four waves per tile;
    kernels of the other workers share the compute units meanwhile, so the
    waves of a block run out of step (an in-place fold must not store before
    every wave of its block holds its operands)."""
    print("server: 4 workers, 12 jobs on two shared params", flush=True)
    cases = []
    for i, (k, log) in enumerate([(8, 12), (32, 11)]):
        A, B, Cm, w = general_instance(log, k, 70 + i)
        p = na.GeneralSnarkParams.secure(log - 5, (A, B, Cm), 0xDEAD, 0xCAFE, k=k)
        cases.append((k, log, p, w, na.prove_general(w, p)))
    srv = na.Server(workers=4)
    ids = [(c, srv.submit_general(c[2], c[3])) for _ in range(6) for c in cases]
    for (k, log, p, w, direct), jid in ids:
        got, _, _ = srv.wait(jid)
        assert got == direct, f"general k={k} log={log}"
    srv.close()


def test_single_proof_beside_a_live_server(orc):
    """While a server exists, single-proof calls post each sumcheck challenge
    before enqueueing the lift that reads it (no kernel waits on a host
    thread); the bytes are the oracle's either way, and a proof made before
    the server and one made while it serves other jobs are identical."""
    print("single proof beside a live server", flush=True)
    A, B, Cm, w = general_instance(10, 32, 81)
    p = na.GeneralSnarkParams.secure(5, (A, B, Cm), 0xDEAD, 0xCAFE, k=32)
    before = na.prove_general(w, p)
    srv = na.Server(workers=2)
    ids = [srv.submit_general(p, w) for _ in range(4)]
    during = na.prove_general(w, p)
    for jid in ids:
        got, _, _ = srv.wait(jid)
        assert got == before
    srv.close()
    assert during == before
    assert before == orc.general_prove(32, 5, A, B, Cm, w)


def test_direct_calls_from_several_threads_match_sequential():
    """Single-proof calls from four threads at once on different params (one
    gets the device to itself and pipelines its rounds, the others wait their
    turn) return the bytes of the same calls made one after another."""
    import threading
    print("direct calls: 4 threads x 3 proofs", flush=True)
    specs = [(32, 12), (8, 11), (64, 10), (16, 12)]
    made = []
    for i, (k, log_n) in enumerate(specs):
        A, B, Cm, w = general_instance(log_n, k, 70 + i)
        made.append((na.GeneralSnarkParams.secure(log_n - 5, (A, B, Cm), 0xDEAD, 0xCAFE, k=k), w))
    want = [na.prove_general(w, p) for p, w in made]
    got = [[None] * 3 for _ in made]
    errs = []

    def work(i):
        try:
            p, w = made[i]
            for r in range(3):
                got[i][r] = na.prove_general(w, p)
        except Exception as e:  # reported below
            errs.append(e)

    th = [threading.Thread(target=work, args=(i,)) for i in range(len(made))]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errs, errs
    for i in range(len(made)):
        assert all(g == want[i] for g in got[i]), f"thread {i}"


def _general(k, log_n, seed):
    A, B, Cm, w = general_instance(log_n, k, seed)
    return na.GeneralSnarkParams.secure(log_n - 5, (A, B, Cm), 0xDEAD, 0xCAFE, k=k), w


def test_job_over_budget_is_refused_at_submit():
    print("server: budget 1 MiB, 2^10 job", flush=True)
    p, w = _general(32, 10, 80)
    srv = na.Server(workers=2, mem_budget=1 << 20)
    with pytest.raises(na.NenyaError) as e:
        srv.submit_general(p, w)
    assert e.value.code == 2
    srv.close()


def test_cancel_queued_and_running_jobs_leaves_the_others_intact():
    import time
    print("server: 1 worker, cancel a queued job and a running one", flush=True)
    big, wb = _general(64, 18, 81)
    small, ws = _general(32, 11, 82)
    want_small = na.prove_general(ws, small)
    srv = na.Server(workers=1)
    a = srv.submit_general(big, wb)
    b = srv.submit_general(small, ws)
    c = srv.submit_general(small, ws)
    srv.cancel(b)  # queued behind the big job on the only worker
    with pytest.raises(na.NenyaError) as e:
        srv.wait(b)
    assert e.value.code == 6
    assert srv.wait(c)[0] == want_small
    assert na.verify_general(big, srv.wait(a)[0]) == "Ok"
    # a running job: the big proof takes long enough to be caught running
    d = srv.submit_general(big, wb)
    time.sleep(0.02)
    try:
        srv.cancel(d)
        canceled = True
    except na.NenyaError as err:  # it had finished already
        assert err.code == 4
        canceled = False
    print(f"running job canceled: {canceled}", flush=True)
    if canceled:
        with pytest.raises(na.NenyaError) as e:
            srv.wait(d)
        assert e.value.code == 6
    else:
        srv.wait(d)
    # the worker goes on proving correctly after the canceled proof
    assert srv.wait(srv.submit_general(small, ws))[0] == want_small
    srv.close()


def test_injected_allocation_failure_fails_only_its_job(monkeypatch):
    print("server: NENYA_INJECT_OOM=2:5, three jobs", flush=True)
    ps = [_general(k, n, 83 + i) for i, (k, n) in enumerate([(32, 11), (16, 12), (8, 11)])]
    want = [na.prove_general(w, p) for p, w in ps]
    monkeypatch.setenv("NENYA_INJECT_OOM", "2:5")
    srv = na.Server(workers=2)
    ids = [srv.submit_general(p, w) for p, w in ps]
    assert srv.wait(ids[0])[0] == want[0]
    with pytest.raises(na.NenyaError) as e:
        srv.wait(ids[1])
    assert e.value.code == 2 and "injected" in str(e.value)
    assert srv.wait(ids[2])[0] == want[2]
    # the same params prove correctly afterwards on the same server
    assert srv.wait(srv.submit_general(*ps[1]))[0] == want[1]
    srv.close()


def test_device_fiat_shamir_server_matches_direct_path(monkeypatch):
    """Workers that draw the round challenges on the device (BLAKE3 absorb
    and squeeze, ChaCha12 sampling and lift in one kernel) produce the bytes
    of the direct path, which runs the transcript on the host."""
    print("server with NENYA_DEVICE_FS=1: general/simd/mimd, all widths", flush=True)
    jobs = []
    for i, (model, k, log) in enumerate([("g", 32, 11), ("g", 64, 10), ("s", 8, 7), ("s", 16, 6), ("m", 64, 6),
                                         ("m", 32, 5), ("g", 8, 12), ("g", 16, 10)]):
        if model == "g":
            A, B, Cm, w = general_instance(log, k, 60 + i)
            p = na.GeneralSnarkParams.secure(log - 5, (A, B, Cm), 0xDEAD, 0xCAFE, k=k)
            jobs.append(("g", p, w, na.prove_general(w, p)))
        elif model == "m":
            As, Bs, Cs, wit = mimd_instance(log, k, 60 + i)
            p = na.MimdSnarkParams.secure(log, log, (As, Bs, Cs), 0xDEAD, 0xCAFE, k=k)
            jobs.append(("m", p, wit, na.prove_mimd(wit, p)))
        else:
            A, B, Cm, wit = simd_instance(log, k, 60 + i)
            p = na.SimdSnarkParams.secure(log, log, 0xDEAD, 0xCAFE, k=k)
            inst = na.SimdR1csInstance(A, B, Cm, wit)
            jobs.append(("s", p, inst, na.prove_simd(inst, p)))
    monkeypatch.setenv("NENYA_DEVICE_FS", "1")
    srv = na.Server(workers=3)
    ids = []
    for model, p, x, _ in jobs:
        ids.append(srv.submit_general(p, x) if model == "g" else srv.submit_mimd(p, x) if model == "m"
                   else srv.submit_simd(x, p))
    for jid, (model, p, x, want) in zip(ids, jobs):
        assert srv.wait(jid)[0] == want, model
    srv.close()

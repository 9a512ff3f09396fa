"""GPU: proofs served with the round hub (NENYA_HUB=1: the small tail rounds
of every running proof's sumchecks fused into shared launches) are
byte-identical to the direct path's, across models, widths and sizes that
hand over at different rounds, and with more concurrent sumchecks than one
fused launch holds."""
import pytest

import nenya_amd as na
from nenya_amd.instances import general_instance, mimd_instance, simd_instance

pytestmark = pytest.mark.gpu


def _jobs(specs, seed0):
    jobs = []
    for i, (model, k, log) in enumerate(specs):
        if model == "g":
            A, B, Cm, w = general_instance(log, k, seed0 + i)
            p = na.GeneralSnarkParams.secure(log - 5, (A, B, Cm), 0xDEAD, 0xCAFE, k=k)
            jobs.append(("g", k, log, p, (A, B, Cm), w))
        elif model == "m":
            As, Bs, Cs, wit = mimd_instance(log, k, seed0 + i)
            p = na.MimdSnarkParams.secure(log, log, (As, Bs, Cs), 0xDEAD, 0xCAFE, k=k)
            jobs.append(("m", k, log, p, (As, Bs, Cs), wit))
        else:
            A, B, Cm, wit = simd_instance(log, k, seed0 + i)
            p = na.SimdSnarkParams.secure(log, log, 0xDEAD, 0xCAFE, k=k)
            jobs.append(("s", k, log, p, (A, B, Cm), wit))
    return jobs


def _submit(srv, job):
    model, k, log, p, m3, w = job
    if model == "g":
        return srv.submit_general(p, w)
    if model == "m":
        return srv.submit_mimd(p, w)
    return srv.submit_simd(na.SimdR1csInstance(*m3, w), p)


def _direct(job):
    model, k, log, p, m3, w = job
    if model == "g":
        return na.prove_general(w, p)
    if model == "m":
        return na.prove_mimd(w, p)
    return na.prove_simd(na.SimdR1csInstance(*m3, w), p)


def test_hub_queue_matches_direct_path(orc, monkeypatch):
    """Six workers over mixed jobs: small ones run every round after the
    first on the hub; SIMD 2^14 and General 2^14 run their first rounds on
    the worker (matrix-core folds) and hand over later."""
    monkeypatch.setenv("NENYA_HUB", "1")
    specs = [("g", 32, 10), ("s", 8, 6), ("m", 64, 6), ("g", 16, 11), ("s", 64, 5), ("s", 32, 14), ("m", 8, 7),
             ("g", 64, 10), ("s", 16, 8), ("g", 32, 14), ("m", 32, 5), ("s", 64, 12), ("g", 8, 12), ("m", 16, 6)]
    print(f"hub server: 6 workers, {len(specs)} mixed jobs", flush=True)
    jobs = _jobs(specs, 500)
    srv = na.Server(workers=6)
    ids = [_submit(srv, j) for j in jobs]
    got = [srv.wait(jid)[0] for jid in ids]
    srv.close()
    for g, job in zip(got, jobs):
        model, k, log, p, m3, w = job
        assert g == _direct(job), f"{model} k={k} log={log}"
    # the oracle's verifier on one of each model
    for g, (model, k, log, p, m3, w) in list(zip(got, jobs))[:3]:
        if model == "g":
            assert orc.general_verify(k, log - 5, *m3, g) == "Ok"
        elif model == "m":
            assert orc.mimd_verify(k, log, log, *m3, g) == "Ok"
        else:
            assert orc.simd_verify(k, log, log, *m3, g) == "Ok"


def test_hub_more_tasks_than_one_launch(monkeypatch):
    """Twenty workers and forty small proofs of one width: more sumchecks are
    on the hub at once than one fused launch holds (16), so a step splits
    its launches."""
    monkeypatch.setenv("NENYA_HUB", "1")
    specs = [("g", 32, 10), ("s", 32, 6), ("m", 32, 6), ("s", 32, 7)] * 10
    print(f"hub server: 20 workers, {len(specs)} u32 jobs", flush=True)
    jobs = _jobs(specs, 900)
    srv = na.Server(workers=20)
    ids = [_submit(srv, j) for j in jobs]
    got = [srv.wait(jid)[0] for jid in ids]
    srv.close()
    for g, job in zip(got, jobs):
        assert g == _direct(job), f"{job[0]} k={job[1]} log={job[2]}"

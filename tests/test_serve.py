"""Served stream: admission policy (CPU) and, on the GPU, every result
identical alone and in the mix (bit-exact), and equal to the plain loop."""

import numpy as np
import pytest

from subsetix_amd.serve import Request, Scheduler


def _req(rid, nbytes):
    return Request(rid, "query", 0.0, {}, nbytes)


def test_scheduler_fifo_budget():
    s = Scheduler(budget_bytes=100)
    for rid, b in enumerate([40, 50, 30, 5]):
        s.offer(_req(rid, b))
    first = s.admit()
    assert [r.rid for r in first] == [0, 1]          # 40 + 50; 30 does not fit
    assert s.admit() == []                            # the head blocks later, smaller requests
    s.release(first[0])
    assert [r.rid for r in s.admit()] == [2, 3]       # 50 + 30 + 5 <= 100
    assert s.used == 85 and s.peak == 90
    with pytest.raises(ValueError):
        s.offer(_req(9, 101))


def test_trace_is_seeded_and_mixed():
    from subsetix_amd.serve import make_trace

    kw = dict(n_fine=(64, 128), steps=(2, 5), query_rows=(16, 64))
    t1 = make_trace(100.0, 40, seed=3, **kw)
    t2 = make_trace(100.0, 40, seed=3, **kw)
    assert [(r.kind, r.arrival) for r in t1] == [(r.kind, r.arrival) for r in t2]
    kinds = [r.kind for r in t1]
    assert 0 < kinds.count("query") < 40
    assert all(np.diff([r.arrival for r in t1]) > 0)
    for r in t1:
        if r.kind == "advect":
            assert r.payload["n_fine"] % 2 == 0 and 2 <= r.payload["steps"] <= 5


def _same(a, b):
    assert a.kind == b.kind and a.rid == b.rid
    if a.kind == "advect":
        assert a.output["mass"] == b.output["mass"]
        np.testing.assert_array_equal(a.output["coarse"], b.output["coarse"])
        np.testing.assert_array_equal(a.output["fine"], b.output["fine"])
    else:
        for op in ("union", "intersection", "difference"):
            for x, y in zip(a.output[op], b.output[op]):
                np.testing.assert_array_equal(x, y)


@pytest.mark.gpu
def test_results_alone_equal_in_mix():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from oracle import oracle as ox
    from subsetix_amd.serve import Server, make_trace, plain_loop

    trace = make_trace(400.0, 10, seed=7, n_fine=(64, 256), steps=(3, 13), query_rows=(32, 300))
    # budget: at most about two advection jobs resident at once -> queueing
    budget = 2 * max(r.est_bytes for r in trace)
    mix = Server(budget_bytes=budget, streams=3, chunk=4, keep_outputs=True).run(trace).results
    assert [r.rid for r in mix] == list(range(10))
    for r in trace:
        print("request", r.rid, r.kind, flush=True)
        alone = Server(budget_bytes=budget, streams=1, chunk=4, keep_outputs=True).run([r]).results[0]
        _same(mix[r.rid], alone)
    plain = plain_loop(trace, keep_outputs=True).results
    for a, b in zip(mix, plain):
        _same(a, b)
    # eager cohort steps (no graphs) on the cohort streams
    eager = Server(budget_bytes=budget, streams=2, chunk=3, use_graph=False, keep_outputs=True).run(trace).results
    for a, b in zip(eager, plain):
        _same(a, b)
    # query results against the oracle
    for r in trace:
        if r.kind != "query":
            continue
        A = ox.make(r.payload["a"]["off"], r.payload["a"]["begin"], r.payload["a"]["end"])
        B = ox.make(r.payload["b"]["off"], r.payload["b"]["begin"], r.payload["b"]["end"])
        for op in ("union", "intersection", "difference"):
            exp = ox.setop(op, A, B)
            off, b, e = mix[r.rid].output[op]
            np.testing.assert_array_equal(off, exp.off)
            np.testing.assert_array_equal(b, exp.beg)
            np.testing.assert_array_equal(e, exp.end)


@pytest.mark.gpu
def test_query_batch_of_several_vs_oracle():
    """Queries stacked into one batch: each query's slice of the stacked
    output against the oracle (the stacked CSR holds no gap rows)."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from oracle import oracle as ox
    from subsetix_amd.serve import query_request, run_queries

    dev = torch.device("cuda", torch.cuda.current_device())
    reqs = [query_request(k, 0.0, rows=r, seed=100 + 7 * k) for k, r in enumerate((300, 17, 1024, 5))]
    outs = run_queries(reqs, dev)
    for r, o in zip(reqs, outs):
        A = ox.make(r.payload["a"]["off"], r.payload["a"]["begin"], r.payload["a"]["end"])
        B = ox.make(r.payload["b"]["off"], r.payload["b"]["begin"], r.payload["b"]["end"])
        for op in ("union", "intersection", "difference"):
            exp = ox.setop(op, A, B)
            off, b, e = o[op]
            np.testing.assert_array_equal(off, exp.off)
            np.testing.assert_array_equal(b, exp.beg)
            np.testing.assert_array_equal(e, exp.end)

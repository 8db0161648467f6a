"""GPU: the request server (continuous batching) returns, for every request,
exactly the stand-alone result, whatever batch it lands in."""
import sys
import threading
import time

import numpy as np
import pytest

from conftest import random_matrix
from gpu_util import assert_same_structure

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _same(a, b):
    assert a.fill_count == b.fill_count
    assert a.te_alg == b.te_alg and a.reach == b.reach
    for x, y in ((a.L_rowptr, b.L_rowptr), (a.L_col, b.L_col), (a.U_rowptr, b.U_rowptr),
                 (a.U_col, b.U_col), (a.row_fill, b.row_fill), (a.super_start, b.super_start),
                 (a.supno, b.supno)):
        np.testing.assert_array_equal(x, y)


def _requests():
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    rng = np.random.default_rng(2024)
    reqs = []
    for k in range(24):
        n = int(rng.integers(1, 400))
        reqs.append(random_matrix(rng, n, float(rng.choice([0.005, 0.02, 0.06])))[0])
    for b, s in enumerate([16, 23, 40, 64, 90]):
        reqs.append(cf.batch_matrix(s, b))
    reqs.append(g.gen_grid(12, 12, 12, nd_leaf=64, seed=7))
    reqs.append(g.gen_band(5000, 256, 8.0, 9))
    return reqs


def test_server_mix_equals_alone(bcsstk18):
    """Requests submitted back to back land in several batches of mixed size;
    each result equals the same matrix run alone (and the oracle for the
    small ones)."""
    import gsofa_amd as g
    from oracle import oracle
    reqs = _requests() + [bcsstk18]
    _say("case: server mix of", len(reqs), "requests, 2 in flight, batches <= 8")
    with g.Server(max_inflight=2, max_batch=8) as srv:
        tickets = [srv.submit(A) for A in reqs]
        out = [srv.wait(t) for t in tickets]
    batches = {info["batch_id"] for _, info in out}
    assert len(batches) >= 4
    assert max(info["batch_size"] for _, info in out) > 1
    for i, (A, (r, info)) in enumerate(zip(reqs, out)):
        _say("check request", i, "n", A.n, "batch", info["batch_id"], "of", info["batch_size"])
        assert info["latency_ms"] >= info["queued_ms"] >= 0.0
        _same(r, g.symbfact(A))
        if A.n <= 2000:
            assert_same_structure(r, oracle.symbfact(A.colbeg, A.colend, A.csr, A.n))


def test_server_concurrent_submitters():
    """Three host threads submit and wait at the same time."""
    import gsofa_amd as g
    reqs = _requests()
    alone = [g.symbfact(A) for A in reqs]
    results = [None] * len(reqs)
    errors = []
    _say("case: server, 3 submitting threads,", len(reqs), "requests")
    with g.Server(max_inflight=3, max_batch=6) as srv:
        def run(idx):
            try:
                ts = [(i, srv.submit(reqs[i])) for i in idx]
                for i, t in ts:
                    results[i] = srv.wait(t)[0]
            except Exception as e:  # surfaced below
                errors.append(e)
        th = [threading.Thread(target=run, args=(list(range(k, len(reqs), 3)),)) for k in range(3)]
        for x in th:
            x.start()
        for x in th:
            x.join()
    assert not errors, errors
    for r, a in zip(results, alone):
        _same(r, a)


def test_server_budget_rejects_oversized():
    """A request that fits the whole budget neither in core nor in the most
    spilled shares (a 200K-row band needs about 80 MB even in 256 shares,
    against 0.05 GiB) fails with GSOFA_ERR_BUDGET; the others complete."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    small = cf.batch_matrix(20, 1)
    big = g.gen_band(200000, 256, 8.0, 3)
    _say("case: server budget 0.05 GiB, one oversized request")
    with g.Server(max_inflight=1, max_batch=4, mem_gb=0.05) as srv:
        t1 = srv.submit(small)
        t2 = srv.submit(big)
        t3 = srv.submit(small)
        r1, _ = srv.wait(t1)
        with pytest.raises(g.GsofaError) as ei:
            srv.wait(t2)
        assert ei.value.code == -7
        r3, _ = srv.wait(t3)
    _same(r1, r3)


def test_server_runs_oversized_alone():
    """A request that does not fit a worker's share but fits the whole budget
    (a 200K-row band: about 1.4 GB, mostly its fill stream and L/U columns
    sized from its row bounds, against 1 GiB shares of 2 GiB) waits for the
    device to drain and runs alone; every result equals its stand-alone run."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    small = cf.batch_matrix(20, 1)
    big = g.gen_band(200000, 256, 8.0, 3)
    _say("case: server budget 2 GiB in 2 shares, one request above a share")
    with g.Server(max_inflight=2, max_batch=4, mem_gb=2.0) as srv:
        ts = [srv.submit(small), srv.submit(big), srv.submit(small), srv.submit(small)]
        out = [srv.wait(t) for t in ts]
    assert out[1][1]["batch_size"] == 1
    _same(out[1][0], g.symbfact(big))
    alone = g.symbfact(small)
    for k in (0, 2, 3):
        _same(out[k][0], alone)


def test_server_spills_what_does_not_fit_in_core():
    """A request whose in-core output does not fit the whole budget (a
    100K-row band of half-width 64 reserves about 0.17 GiB for its per-row
    arrays, fill stream and L/U columns, sized from its exact row bounds,
    against 0.15 GiB) runs alone in spilled shares and equals its in-core
    stand-alone run."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    small = cf.batch_matrix(20, 1)
    big = g.gen_band(100000, 64, 8.0, 3)
    assert g.footprint(big) > int(0.15 * 2**30)
    _say("case: server budget 0.15 GiB, one request served in spilled shares")
    with g.Server(max_inflight=1, max_batch=4, mem_gb=0.15) as srv:
        ts = [srv.submit(small), srv.submit(big), srv.submit(small)]
        out = [srv.wait(t) for t in ts]
    assert out[1][1]["spill_shares"] > 1 and out[1][1]["batch_size"] == 1
    assert out[0][1]["spill_shares"] == 1
    _same(out[1][0], g.symbfact(big))
    _same(out[0][0], out[2][0])


def test_server_poll_and_unknown_ticket():
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    A = cf.batch_matrix(30, 2)
    with g.Server() as srv:
        t = srv.submit(A)
        got = None
        for _ in range(2000):  # up to ~10 s of polling, then block
            got = srv.poll(t)
            if got is not None:
                break
            time.sleep(0.005)
        if got is None:
            got = srv.wait(t)
        _same(got[0], g.symbfact(A))
        with pytest.raises(g.GsofaError):
            srv.poll(t)  # taken


def test_server_malformed_request_is_refused_alone():
    """A request with a column index out of range is refused at submit, to its
    caller alone; the valid requests around it complete unaffected."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    good = [cf.batch_matrix(20, 1), cf.batch_matrix(33, 2)]
    bad = g.CsrMatrix.from_rows([[0, 1], [0, 1, 5], [1, 2]])  # 5 >= n = 3
    neg = g.CsrMatrix.from_rows([[0, -1], [1]])
    _say("case: server, malformed requests between valid ones")
    with g.Server(max_inflight=1, max_batch=8) as srv:
        t0 = srv.submit(good[0])
        for B in (bad, neg):
            with pytest.raises(g.GsofaError) as ei:
                srv.submit(B)
            assert ei.value.code == -1
        t1 = srv.submit(good[1])
        out = [srv.wait(t0)[0], srv.wait(t1)[0]]
    for A, r in zip(good, out):
        _same(r, g.symbfact(A))


def test_server_two_waiters_on_one_ticket():
    """Two threads wait on the same ticket: one takes the result, the other
    gets GSOFA_ERR_INVALID (the ticket is gone), and nothing is freed twice."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    A = cf.batch_matrix(60, 3)
    got, errs = [], []
    _say("case: server, two waiters on one ticket")
    with g.Server(max_inflight=1, max_batch=4) as srv:
        t = srv.submit(A)

        def waiter():
            try:
                got.append(srv.wait(t)[0])
            except g.GsofaError as e:
                errs.append(e.code)
        th = [threading.Thread(target=waiter) for _ in range(2)]
        for x in th:
            x.start()
        for x in th:
            x.join()
    assert len(got) == 1 and errs == [-1]
    _same(got[0], g.symbfact(A))


def test_server_counts_only_matches_oracle():
    """A counts-only server returns each request's fill count, per-row fills,
    TE and |R| (no L/U), equal to the oracle's (per-row TE and |R| where the
    mode reports them: in reach mode their totals)."""
    import gsofa_amd as g
    from oracle import oracle
    reqs = _requests()[:20]
    _say("case: counts-only server,", len(reqs), "requests")
    with g.Server(max_inflight=2, max_batch=6, counts_only=True) as srv:
        ts = [srv.submit(A) for A in reqs]
        out = [srv.wait(t)[0] for t in ts]
    for A, r in zip(reqs, out):
        o = oracle.symbfact(A.colbeg, A.colend, A.csr, A.n)
        assert r.fill_count == o.fill_count and r.te_alg == o.te_alg and r.reach == o.reach
        np.testing.assert_array_equal(r.row_fill, o.row_fill)
        if r.row_te is not None:
            np.testing.assert_array_equal(r.row_te, o.row_te)
            np.testing.assert_array_equal(r.row_reach, o.row_reach)
        assert r.L_rowptr is None and r.n_super == -1

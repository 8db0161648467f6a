"""GPU: the request stream of the serving workload (Poisson arrivals of 2D
nested-dissection grids, side 16..256, and 3D 7-point nested-dissection
grids, side 16..64) submitted to the server at its arrival times by
concurrent clients.  Every result equals the oracle's golden for its matrix
(totals, L/U digest, per-row fills, supernodes), whatever batch it landed in
and whatever ran beside it."""
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import pytest

from conftest import ROOT
from digests import array_digest
from gpu_util import record_of

pytestmark = pytest.mark.gpu

GOLD = os.path.join(ROOT, "tests", "golden", "goldens.json")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("rate", [200.0, 2000.0])
def test_served_stream_matches_goldens(rate):
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    if not os.path.exists(GOLD):
        pytest.skip("goldens not generated")
    golds = json.load(open(GOLD)).get("stream")
    if not golds:
        pytest.skip("stream goldens not generated")
    reqs = cf.stream(len(golds), rate)
    assert [(k, s) for _, k, s, _ in reqs] == [(x["kind"], x["side"]) for x in golds]
    assert max(s for _, k, s, _ in reqs if k == 1) >= 48
    _say("case: served stream of", len(reqs), "requests at", rate, "req/s")
    with g.Server(max_inflight=4, max_batch=64) as srv:
        t0 = time.perf_counter()
        futs = []
        with ThreadPoolExecutor(max_workers=8) as pool:
            for ta, _, _, A in reqs:
                now = time.perf_counter() - t0
                if ta > now:
                    time.sleep(ta - now)
                futs.append(pool.submit(srv.submit, A))
            tickets = [f.result() for f in futs]
        out = [srv.wait(t) for t in tickets]
    batches = {info["batch_id"] for _, info in out}
    _say("batches", len(batches), "largest", max(info["batch_size"] for _, info in out))
    assert len(batches) >= 2 and max(info["batch_size"] for _, info in out) > 1
    for i, ((r, info), gold) in enumerate(zip(out, golds)):
        _say("check request", i, "kind", gold["kind"], "side", gold["side"], "batch", info["batch_id"],
             "of", info["batch_size"])
        rec = record_of(r)
        for k in ("fill_count", "te_alg", "reach", "nnz_L", "nnz_U", "lu", "row_fill"):
            assert rec[k] == gold[k], (i, k)
        assert r.n_super == gold["n_super0"], i
        assert array_digest(r.super_start) == gold["super0"], i
        assert info["latency_ms"] >= info["queued_ms"] >= 0.0

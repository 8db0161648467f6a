"""GPU: the label re-initialisation that the decreasing max-id offset needs
every 2^32 / n groups (with_space_optimization/symbfact.cuh:1928-1932,
2388-2407).  At real sizes a plan reaches it only after many executes (about
80 at n = 659,033), so GSOFA_OFFSET_SPAN, read at plan creation, lowers the
span to a few groups here; every execute must stay bit-exact in each mode.
Reach mode has no labels, but its visited words carry a per-workgroup stamp
that the same span bounds: the words are cleared when the stamps run out."""
import sys

import pytest

from gpu_util import assert_same_structure

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [
    ("block 32 slots", dict(mode="block")),
    ("block 8 slots", dict(mode="block", block_slots=8)),
    ("grid, chunk 1024", dict(mode="grid", chunk=1024)),
    ("hybrid, top 2048", dict(mode="block", top_sources=2048)),
    ("reach 64 slots", dict(mode="reach", block_slots=64)),
    ("reach 32 slots", dict(mode="reach", block_slots=32)),
]


@pytest.mark.parametrize("span", [1, 2, 5])
@pytest.mark.parametrize("name,kw", CASES)
def test_offset_reinit_repeated(bcsstk18, monkeypatch, span, name, kw):
    import gsofa_amd as g
    from oracle import oracle
    o = oracle.symbfact(bcsstk18.colbeg, bcsstk18.colend, bcsstk18.csr, bcsstk18.n)
    monkeypatch.setenv("GSOFA_OFFSET_SPAN", str(span))
    _say("case: offset span", span, name)
    with g.Plan(bcsstk18, **kw) as p:
        for it in range(3):
            p.execute()
            assert_same_structure(p.fetch(), o)


def test_offset_reinit_batch(monkeypatch):
    """A batch re-initialises each workgroup's labels between every two
    mini-groups; every matrix still equals the oracle."""
    import numpy as np
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    from conftest import random_matrix
    from oracle import oracle
    rng = np.random.default_rng(5)
    mats = [cf.batch_matrix(s, b) for b, s in enumerate([16, 40, 77])]
    mats += [random_matrix(rng, int(rng.integers(2, 300)), 0.02)[0] for _ in range(5)]
    monkeypatch.setenv("GSOFA_OFFSET_SPAN", "2")
    _say("case: offset span 2, batch of", len(mats))
    with g.Plan(mats) as p:
        for it in range(2):
            p.execute()
            for b, A in enumerate(mats):
                assert_same_structure(p.fetch_matrix(b), oracle.symbfact(A.colbeg, A.colend, A.csr, A.n))

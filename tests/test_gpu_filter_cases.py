"""The staged filter kernel (bms2_filter_candidates) on the hand-derived iteration-0 domains of
tests/filter_cases.py: one case per filter term (label or wildcard, degree, path containment, ring
counts), each excluding candidates that the other three terms keep.  The fused pipeline, which
screens with the same four terms, must give the GMCR the staged path gives on the same inputs."""
import numpy as np
import pytest

import filter_cases
from bms2_amd import api

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def queue():
    q = api.Queue(0)
    yield q
    q.close()


@pytest.mark.parametrize("case", filter_cases.CASES, ids=[c[0] for c in filter_cases.CASES])
def test_filter_kernel_iteration0_domains(queue, case):
    qb, db = filter_cases.batches(case)
    print(f"filter case {case[0]}: {qb.total_nodes} query nodes, {db.total_nodes} data nodes",
          flush=True)
    out = api.run_staged(queue, qb, db, K=0, keep_domains=True)
    got = filter_cases.domain_sets(out["domains"][0], qb.total_nodes, db.total_nodes)
    assert got == case[3]
    # GMCR: the query is valid iff every domain is non-empty (it has more than one node)
    off, idx = out["gmcr"]
    expect_valid = all(case[3])
    assert (len(idx) == 1) == expect_valid


def test_filter_cases_in_one_batch(queue):
    """All cases' graphs in one query batch against all data graphs in one batch: domains are
    rows over global data nodes, so inside its own case's data graph each query node's domain is
    the expected set shifted by that graph's first node."""
    from bms2_amd.graphs import GraphBatch
    qb = GraphBatch.from_graphs([c[1] for c in filter_cases.CASES])
    db = GraphBatch.from_graphs([c[2] for c in filter_cases.CASES])
    out = api.run_staged(queue, qb, db, K=0, keep_domains=True)
    got = filter_cases.domain_sets(out["domains"][0], qb.total_nodes, db.total_nodes)
    row = 0
    for i, c in enumerate(filter_cases.CASES):
        base = int(db.graph_offsets[i])
        for qn_expect in c[3]:
            inside = {d - base for d in got[row]
                      if base <= d < int(db.graph_offsets[i + 1])}
            assert inside == qn_expect, (c[0], row)
            row += 1
    # the fused pipeline's GMCR bits equal the staged GMCR
    counts, valid, _ = api.run_pipeline(queue, qb, db, 0, False)
    off, idx = out["gmcr"]
    bits = np.unpackbits(valid.view(np.uint8), axis=1, bitorder="little")[:, :qb.num_graphs]
    for g in range(db.num_graphs):
        assert set(np.flatnonzero(bits[g])) == set(idx[off[g]:off[g + 1]].tolist()), g

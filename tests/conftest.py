"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on the CPU."""
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def molecules_2k():
    """2,000 molecules of 10-40 atoms x 64 queries of 4-10 nodes (seed 0)."""
    from bms2_amd import graphs
    return graphs.generate("molecules-2k", seed=0)


@pytest.fixture(scope="session")
def molecules_1m_shard():
    """First 3,000 graphs of the 1M-molecule workload with its 1,024 queries (seed 0)."""
    from bms2_amd import graphs
    return graphs.generate("molecules-1m", seed=0, num_data_graphs=3000)


def expected_stats(queries, data, K, find_first, rs, dedup=True):
    """The statistics bms2_pipeline_run reports next to the oracle's `rs` over every query: GMCR
    pairs, matched pairs and embeddings are those of every query; DFS nodes are those of the
    queries the run screens, one per isomorphism class when dedup is on (the others copy their
    class's results), i.e. the oracle's DFS nodes over the classes' first queries."""
    import numpy as np

    import oracle
    from bms2_amd import api, graphs
    if not dedup:
        return rs
    rep = api.query_classes(queries)
    keep = np.flatnonzero(rep == np.arange(len(rep)))
    if len(keep) == queries.num_graphs:
        return rs
    sub = graphs.GraphBatch.from_graphs([queries.graph(int(i)) for i in keep])
    out = np.array(rs, copy=True)
    out[3] = oracle.pipeline(sub, data, K, find_first)[2][3]
    return out

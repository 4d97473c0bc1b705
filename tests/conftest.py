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
    pairs, matched pairs and embeddings are those of every query; with dedup on, DFS nodes are
    those of the queries the run screens, one per isomorphism class (the others copy their
    class's results), with the symmetry breaking of bms2_query_symmetry in the count-all searches
    on data graphs of up to 64 nodes: the oracle's staged join over the classes' first queries
    under those constraints."""
    import numpy as np

    import oracle
    from bms2_amd import api, graphs
    if not dedup:
        return rs
    rep = api.query_classes(queries)
    keep = np.flatnonzero(rep == np.arange(len(rep)))
    sub = queries if len(keep) == queries.num_graphs else graphs.GraphBatch.from_graphs(
        [queries.graph(int(i)) for i in keep])
    du, vm = api.query_symmetry(sub)
    if find_first:  # the symmetry breaking serves count-all mode only
        du[:], vm[:] = 0xff, 0
    if len(keep) == queries.num_graphs and not vm.any():
        return rs
    r = oracle.staged(sub, data, K, find_first, keep_domains=False)
    _, _, nodes = oracle.join(sub, data, r.final_domains, r.gmcr_offsets, r.gmcr_indices,
                              find_first, None, sym=(du, vm))
    # graphs above 64 nodes are joined by the big-graph path, without the symmetry breaking
    sizes = np.diff(np.asarray(data.graph_offsets, np.int64))
    pair_g = np.repeat(np.arange(data.num_graphs), np.diff(r.gmcr_offsets.astype(np.int64)))
    big = sizes[pair_g] > 64
    out = np.array(rs, copy=True)
    out[3] = int(nodes[~big].sum()) + int(r.dfs_nodes[big].sum())
    return out

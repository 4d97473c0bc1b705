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

"""The C++ driver built on include/sigmo.hpp (the reference's BMS2 stage order) on .dat input."""
import re
import subprocess
from pathlib import Path

import pytest

import oracle
from bms2_amd import graphs

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parents[1] / "bms2_amd" / "bms2_main"


@pytest.mark.parametrize("K,find_all", [(0, True), (2, True), (6, True), (6, False)])
def test_driver_matches_oracle(tmp_path, molecules_2k, K, find_all):
    data, queries = molecules_2k
    sub = data.subset(0, 400)
    dpath, qpath = tmp_path / "data.dat", tmp_path / "query.dat"
    graphs.write_dat(str(dpath), sub)
    graphs.write_dat(str(qpath), queries)
    cmd = [str(EXE), "-Q", str(qpath), "-D", str(dpath), "-i", str(K)]
    if find_all:
        cmd.append("--find-all")
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    m = re.search(r"# Matches: (\d+)", out.stdout)
    assert m, out.stdout
    # the oracle on the same .dat content (reloaded, so CSR neighbour order matches the file)
    d2, q2 = graphs.load_dat(str(dpath)), graphs.load_dat(str(qpath))
    ref = oracle.staged(q2, d2, K, find_first=not find_all, keep_domains=False)
    assert int(m.group(1)) == ref.total


@pytest.mark.parametrize("K,find_all", [(2, True), (6, True), (6, False)])
def test_driver_use_cs_matches_oracle(tmp_path, molecules_2k, K, find_all):
    """bms2_main --use-cs on a cluster file (prototype first): the CS join's total."""
    data, _ = molecules_2k
    sub = data.subset(0, 400)
    dpath, qpath = tmp_path / "data.dat", tmp_path / "cluster.dat"
    graphs.write_dat(str(dpath), sub)
    graphs.write_dat(str(qpath), graphs.cluster_queries(sub, 15, seed=K))
    cmd = [str(EXE), "-Q", str(qpath), "-D", str(dpath), "-i", str(K), "--use-cs"]
    if find_all:
        cmd.append("--find-all")
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    m = re.search(r"# Matches: (\d+)", out.stdout)
    assert m and "partial_match_depth set to" in out.stdout, out.stdout
    d2, q2 = graphs.load_dat(str(dpath)), graphs.load_dat(str(qpath))
    ref = oracle.staged(q2, d2, K, find_first=not find_all, keep_domains=False)
    end = int(q2.graph_offsets[1] - q2.graph_offsets[0])
    part = oracle.join_partial(q2, d2, ref.final_domains, end)
    total, _ = oracle.join_with_partial(q2, d2, ref.final_domains, ref.gmcr_offsets,
                                        ref.gmcr_indices, end, part, not find_all)
    assert int(m.group(1)) == total
    m2 = re.search(r"Found (\d+) partial matches", out.stdout)
    assert m2 and int(m2.group(1)) == len(part)

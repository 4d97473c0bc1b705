"""The C++ driver built on include/sigmo.hpp (the reference's BMS2 stage order) on .dat input."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from bms2_amd import graphs

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parents[1] / "bms2_amd" / "bms2_main"
BATCH = EXE.with_name("bms2_batch")


def _stat(stdout: str, name: str) -> int:
    """A '# name: N' result line; numbers carry '.' every three digits (formatNumber)."""
    m = re.search(rf"# {name}: ([\d.]+)", stdout)
    assert m, stdout
    return int(m.group(1).replace(".", ""))


def _matches(stdout: str) -> int:
    return _stat(stdout, "Matches")


def _run(args, timeout=120):
    out = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=timeout)
    assert out.returncode == 0, out.stderr
    return out.stdout


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
    # the oracle on the same .dat content (reloaded, so CSR neighbour order matches the file)
    d2, q2 = graphs.load_dat(str(dpath)), graphs.load_dat(str(qpath))
    ref = oracle.staged(q2, d2, K, find_first=not find_all, keep_domains=False)
    assert _matches(out.stdout) == ref.total
    # candidate statistics of BMS2.cpp:373-385 from the final domains
    counts = np.unpackbits(ref.final_domains.view(np.uint8), axis=1,
                           bitorder="little").sum(axis=1)
    assert _stat(out.stdout, "Total candidates") == int(counts.sum())
    assert _stat(out.stdout, "Zero candidates") == int((counts == 0).sum())
    assert _stat(out.stdout, "Median candidates") == int(np.sort(counts)[len(counts) // 2])


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
    assert "partial_match_depth set to" in out.stdout, out.stdout
    d2, q2 = graphs.load_dat(str(dpath)), graphs.load_dat(str(qpath))
    ref = oracle.staged(q2, d2, K, find_first=not find_all, keep_domains=False)
    end = int(q2.graph_offsets[1] - q2.graph_offsets[0])
    part = oracle.join_partial(q2, d2, ref.final_domains, end)
    total, _ = oracle.join_with_partial(q2, d2, ref.final_domains, ref.gmcr_offsets,
                                        ref.gmcr_indices, end, part, not find_all)
    assert _matches(out.stdout) == total
    m2 = re.search(r"Found ([\d.]+) partial matches", out.stdout)
    assert m2 and int(m2.group(1).replace(".", "")) == len(part)


@pytest.fixture()
def dat_files(tmp_path, molecules_2k):
    data, queries = molecules_2k
    dpath, qpath = tmp_path / "data.dat", tmp_path / "query.dat"
    graphs.write_dat(str(dpath), data.subset(0, 600))
    graphs.write_dat(str(qpath), queries)
    return dpath, qpath, graphs.load_dat(str(dpath)), graphs.load_dat(str(qpath))


def test_driver_flags(dat_files):
    dpath, qpath, d2, q2 = dat_files
    base = ["-Q", qpath, "-D", dpath, "-i", 2, "--find-all"]
    ref = oracle.staged(q2, d2, 2, keep_domains=False)
    # the fused pipeline gives the stage calls' total
    out = _run([EXE, *base, "--fused"])
    assert _matches(out) == ref.total and "Fused pipeline: Yes" in out
    # -m 2: queries and data graphs twice each, four times the matches
    assert _matches(_run([EXE, *base, "-m", "2"])) == 4 * ref.total
    # --query-filter keeps queries of 5-7 nodes
    sizes = np.diff(q2.graph_offsets)
    keep = [i for i in range(q2.num_graphs) if 5 <= sizes[i] <= 7]
    qf = graphs.GraphBatch.from_graphs([q2.graph(i) for i in keep])
    reff = oracle.staged(qf, d2, 2, keep_domains=False)
    assert _matches(_run([EXE, *base, "--query-filter", "5:7"])) == reff.total
    # --skip-join: no match line; --skip-candidates-analysis: no candidate lines
    out = _run([EXE, *base, "--skip-join", "--skip-candidates-analysis"])
    assert "# Matches" not in out and "# Total candidates" not in out and "Join time: skipped" in out
    # bad input is a usage error
    bad = subprocess.run([str(EXE), "-Q", str(qpath)], capture_output=True, text=True)
    assert bad.returncode == 2 and "Both query and data files" in bad.stderr


@pytest.mark.parametrize("fused", [False, True])
def test_batch_driver(dat_files, fused):
    """bms2_batch streams the data file in batches; the total equals the whole file's."""
    dpath, qpath, d2, q2 = dat_files
    extra = ["--fused"] if fused else []
    out = _run([BATCH, "-Q", qpath, "-D", dpath, "-i", 2, "--find-all", "--batch-size", 250, *extra])
    assert out.count(">>> Processing Batch") == 3  # 600 graphs in batches of 250
    assert "Total Data Graphs: 600" in out
    ref = oracle.staged(q2, d2, 2, keep_domains=False)
    assert _matches(out) == ref.total
    per_batch = [int(x) for x in re.findall(r"-> Matches found: (\d+)", out)]
    assert sum(per_batch) == ref.total
    # first-match mode stops after the first batch with a match (BMS2_batch.cpp:307-310)
    out = _run([BATCH, "-Q", qpath, "-D", dpath, "-i", 2, "--batch-size", 250, *extra])
    assert "Stopping." in out and out.count(">>> Processing Batch") == 1
    r1 = oracle.staged(q2, d2.subset(0, 250), 2, find_first=True, keep_domains=False)
    assert _matches(out) == r1.total
    # --max-data-graphs cuts the stream
    out = _run([BATCH, "-Q", qpath, "-D", dpath, "-i", 2, "--find-all", "--batch-size", 250,
                "--max-data-graphs", 300, *extra])
    assert "Total Data Graphs: 300" in out
    r3 = oracle.staged(q2, d2.subset(0, 300), 2, keep_domains=False)
    assert _matches(out) == r3.total

"""The C++ facade (include/sigmo.hpp) as a drop-in for the reference's driver calls.

tests/dropin/bms2_callshapes.cpp makes the library calls of library/src/BMS2.cpp in the same
shapes (sigmo:: in place of sycl::): the allocation-size getters, DeviceOptions, device info,
malloc_shared / memset / free, the CS join without a capacity argument, and the host candidates'
getCandidatesCount(i).  The CPU tests compile it with warnings as errors; the GPU tests run the
built driver on .dat files and compare its report with the oracle."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "dropin" / "bms2_callshapes.cpp"
EXE = SRC.with_suffix("")


def test_reference_call_shapes_compile():
    out = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                          "-I", str(ROOT / "include"), str(SRC)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr


def test_facade_compiles_alone_pedantic(tmp_path):
    src = tmp_path / "one.cpp"
    src.write_text('#include <sigmo.hpp>\nint main() { return 0; }\n')
    out = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Wpedantic", "-Werror",
                          "-fsyntax-only", "-I", str(ROOT / "include"), str(src)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr


def test_callshapes_driver_is_built():
    assert EXE.exists(), "build() compiles tests/dropin/bms2_callshapes"


def _stat(stdout: str, name: str) -> int:
    m = re.search(rf"# {name}: (\d+)", stdout)
    assert m, stdout
    return int(m.group(1))


@pytest.mark.gpu
@pytest.mark.parametrize("K,find_all", [(0, True), (6, True), (6, False)])
def test_callshapes_driver_matches_oracle(tmp_path, molecules_2k, K, find_all):
    import oracle
    from bms2_amd import graphs
    data, queries = molecules_2k
    dpath, qpath = tmp_path / "data.dat", tmp_path / "query.dat"
    graphs.write_dat(str(dpath), data.subset(0, 300))
    graphs.write_dat(str(qpath), queries)
    print(f"callshapes K={K} find_all={find_all}", flush=True)
    out = subprocess.run([str(EXE), str(qpath), str(dpath), str(K), str(int(find_all)), "0"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    d2, q2 = graphs.load_dat(str(dpath)), graphs.load_dat(str(qpath))
    ref = oracle.staged(q2, d2, K, find_first=not find_all, keep_domains=False)
    assert _stat(out.stdout, "Matches") == ref.total
    counts = np.unpackbits(ref.final_domains.view(np.uint8), axis=1,
                           bitorder="little").sum(axis=1)
    assert _stat(out.stdout, "Total candidates") == int(counts.sum())
    assert _stat(out.stdout, "Zero candidates") == int((counts == 0).sum())
    # allocation sizes: signatures 40 + path 32 + cycle 8 bytes per node, candidates u64 words
    m = re.search(r"Total allocated memory: (\d+) out of (\d+)", out.stdout)
    assert m and int(m.group(2)) > 100 * 2**30


@pytest.mark.gpu
@pytest.mark.parametrize("find_all", [True, False])
def test_callshapes_driver_cs_matches_oracle(tmp_path, molecules_2k, find_all):
    """The CS path as BMS2.cpp calls it: joinPartialCandidates2 with no capacity (taken from the
    malloc_shared buffer), joinWithPartialMatches with pair_done flags in shared memory."""
    import oracle
    from bms2_amd import graphs
    data, _ = molecules_2k
    sub = data.subset(0, 300)
    dpath, qpath = tmp_path / "data.dat", tmp_path / "cluster.dat"
    graphs.write_dat(str(dpath), sub)
    graphs.write_dat(str(qpath), graphs.cluster_queries(sub, 15, seed=2))
    print(f"callshapes CS find_all={find_all}", flush=True)
    out = subprocess.run([str(EXE), str(qpath), str(dpath), "2", str(int(find_all)), "1"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    d2, q2 = graphs.load_dat(str(dpath)), graphs.load_dat(str(qpath))
    ref = oracle.staged(q2, d2, 2, find_first=not find_all, keep_domains=False)
    end = int(q2.graph_offsets[1] - q2.graph_offsets[0])
    part = oracle.join_partial(q2, d2, ref.final_domains, end)
    total, per = oracle.join_with_partial(q2, d2, ref.final_domains, ref.gmcr_offsets,
                                          ref.gmcr_indices, end, part, not find_all)
    assert _stat(out.stdout, "Matches") == total
    assert f"Found {len(part)} partial matches" in out.stdout
    if not find_all:
        assert _stat(out.stdout, "Pairs done") == int((per > 0).sum())

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

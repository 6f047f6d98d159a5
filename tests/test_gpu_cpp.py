"""The reference's u16_centrality acceptance tests, restated in C++ against the
C++ mirror of the call surface (tests/cpp/u16_centrality.cpp), run on the GPU."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "u16_centrality.bin"

pytestmark = pytest.mark.gpu


def test_u16_centrality_in_cpp(gpu):
    assert BIN.exists(), "build it with __graft_entry__.build()"
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout

"""The C++ drop-in test (tests/cpp/test_dropin.cpp) on the GPU: DiVerG's call
sites compiled unchanged against diverg_amd/include, checked against the
oracle."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_dropin")


def test_cpp_dropin():
    assert os.path.exists(BIN), "build it with __graft_entry__.build()"
    gfa = os.path.join(HERE, "golden", "gfa", "multi.gfa")
    res = subprocess.run([BIN, gfa], capture_output=True, text=True, timeout=100)
    print(res.stdout[-2000:], res.stderr[-2000:], flush=True)
    assert res.returncode == 0
    assert "passed (0 failures)" in res.stdout

"""Runs the C++ drop-in layer tests (tests/cpp/test_dropin.cpp) on the GPU."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "test_dropin"


@pytest.mark.gpu
def test_cpp_dropin_layer():
    if not BIN.exists():
        pytest.fail(f"{BIN} not built (run __graft_entry__.build())")
    print("[case] C++ drop-in layer tests", flush=True)
    res = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=300)
    print(res.stdout)
    print(res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr

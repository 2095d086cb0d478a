"""Runs the C++ parity tests of include/sil/exterior.hpp (tests/cpp) on the GPU."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).with_name("cpp") / "exterior_tests"


def test_cpp_exterior_suite():
    assert BIN.exists(), "build with `python -m similie_amd.build`"
    p = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr

"""Runs tests/cpp/dropin_test (built by __graft_entry__.build()): the C++
drop-in headers used with the reference's class and function signatures."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "dropin_test")


def test_cpp_dropin_program(tmp_path):
    """The program writes its tree file and raw volume in its working
    directory, here a temporary one."""
    assert os.path.exists(BIN), "tests/cpp/_build/dropin_test is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    print(r.stdout[-2000:])
    assert r.returncode == 0 and "DROPIN OK" in r.stdout

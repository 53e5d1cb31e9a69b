"""Construction's subtree minima and ID arrays (filled during construction)
against the same arrays recomputed level by level from the built nodes,
through tests/cpp/submin_test.cpp."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "submin_test")


def test_construction_subtree_minima_equal_levelwise():
    assert os.path.exists(BIN), "tests/cpp/_build/submin_test is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "all subtree minima ok" in r.stdout

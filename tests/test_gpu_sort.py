"""The library's radix sorts (construction, level index and query grouping
sort with them) against std::stable_sort, through tests/cpp/sort_test.cpp."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "sort_test")


def test_radix_sorts_equal_stable_sort():
    assert os.path.exists(BIN), "tests/cpp/_build/sort_test is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "all sorts ok" in r.stdout

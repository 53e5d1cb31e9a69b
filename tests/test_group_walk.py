"""Runs tests/cpp/group_walk_test (built by __graft_entry__.build()): the
grouped walk's per-node transition (node_step), walked on the CPU over
oracle-built trees, must give every query exactly the oracle's report set."""
import os
import subprocess

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "group_walk_test")


def test_group_walk_matches_oracle():
    assert os.path.exists(BIN), "tests/cpp/_build/group_walk_test is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0 and "GROUP WALK OK" in r.stdout

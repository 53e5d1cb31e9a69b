"""The vectorised whole-tree checks used on 2^26-point GPU trees, pinned on
the CPU: they accept the oracle's trees (themselves pinned to the
reference's), agree with the oracle's closed-form shape, and reject trees
with a planted mistake of each kind they are meant to catch."""
import numpy as np
import pytest

import oracle as O
import tree_invariants as TI


@pytest.mark.parametrize("n", [1, 2, 3, 7, 8, 100, 1000, 4097, 30001])
def test_shape_matches_oracle(n):
    assert np.array_equal(TI.shape_sizes(n), O.shape_sizes(n))


def _tree(n, seed):
    import pst_data as D
    pts = D.interval_points(n, seed=seed)
    return pts, O.build(pts, "f32", "u32").arrays()


@pytest.mark.parametrize("n,seed", [(1, 0), (5, 1), (1000, 2), (65537, 3)])
def test_oracle_trees_pass(n, seed):
    pts, ex = _tree(n, seed)
    TI.check(ex, pts)


def _planted(ex, what, rng):
    ex = {k: v.copy() for k, v in ex.items()}
    S = len(ex["dim1"])
    used = np.flatnonzero(ex["bits"] != 0)  # slots with a child
    s = int(rng.choice(used))
    c = 2 * s + 1
    if what == "swap_parent_child":      # heap order
        for k in ("dim1", "dim2", "ids"):
            ex[k][[s, c]] = ex[k][[c, s]]
    elif what == "med_off_by_one_point":  # med taken from the wrong point
        ex["med"][s] = np.nextafter(ex["med"][s], np.float32(2))
    elif what == "swap_siblings":        # BST split
        if ex["bits"][s] == 3:
            for k in ("dim1", "dim2", "ids", "med"):
                ex[k][[c, c + 1]] = ex[k][[c + 1, c]]
        else:
            ex["bits"][s] ^= 1
    elif what == "bitcode":
        ex["bits"][s] ^= 2
    elif what == "dup_id":
        ex["ids"][c] = ex["ids"][s]
    elif what == "dim2_changed":
        ex["dim2"][c] = np.nextafter(ex["dim2"][c], np.float32(-1))
    return ex


@pytest.mark.parametrize("what", ["swap_parent_child", "med_off_by_one_point", "swap_siblings", "bitcode",
                                  "dup_id", "dim2_changed"])
def test_planted_mistakes_fail(what):
    pts, ex = _tree(5000, 7)
    rng = np.random.default_rng(hash(what) & 0xffff)
    bad = _planted(ex, what, rng)
    with pytest.raises(AssertionError):
        TI.check(bad, pts)

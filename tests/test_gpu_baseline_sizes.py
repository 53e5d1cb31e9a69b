"""Parity at the workloads' full sizes: the 2^26-point float32 tree of the
plain GPU workload and the 2^27-point int32 tree of the batched workload.

- The 2^26 tree passes every whole-tree invariant (closed-form shape and
  bitcodes, ID permutation, heap order on (dim2, id), BST split and med on
  (dim1, id)), which together fix the reference's tree.
- A batch of the workload's queries with the generator's half-left /
  half-right two-sided mix (right (q, q) queries report about n/2 IDs) and
  three-sided queries runs through the tree walk and through the level
  index: every count equals the oracle's walk on the exported tree (the
  oracle is pinned to the reference's StaticPSTCPUIter), and the full ID sets
  of 16 sampled queries equal a vectorised brute force.
- The int32 batch of 8192 mixed queries runs as a captured HIP graph after
  autotuning, through the walk and through the level index, with the same
  checks.
"""
import os

import numpy as np
import pytest

import oracle as O
import tree_invariants as TI

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

THREADS = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 8))


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


@pytest.fixture(scope="module")
def P():
    import pst_amd
    pst_amd.lib()
    return pst_amd


def _brute(pts, q):
    kind = int(q["kind"])
    m = pts["dim2"] >= q["min_dim2"]
    if kind != 2:
        m &= pts["dim1"] <= q["max_dim1"]
    if kind != 1:
        m &= pts["dim1"] >= q["min_dim1"]
    return np.flatnonzero(m).astype(np.uint32)


def _oracle_tree(P, tree, t, n):
    import ctypes
    tr = O.empty_tree(t, "u32", n)
    nb = P.lib().pst_gpu_export_bytes(tree.h)
    off = (ctypes.c_size_t * 6)()
    P._check(P.lib().pst_gpu_export_tree(tree.h, tr.blob.ctypes.data, nb, off))
    return tr


def _run_batch(torch, P, tree, qs, total, graph=False):
    nq = len(qs)
    dq = P.to_device(qs)
    counts = torch.zeros(nq, dtype=torch.int64, device="cuda")
    offs = torch.zeros(nq + 1, dtype=torch.int64, device="cuda")
    out = torch.empty(total + 1, dtype=torch.int32, device="cuda")
    if graph:
        s = torch.cuda.Stream()
        g = tree.batch_graph(dq, nq, counts, offs, out, total, stream=s)
        for _ in range(2):
            with torch.cuda.stream(s):
                counts.zero_(), offs.zero_()
            g.launch(s)
            s.synchronize()
        g.close()
    else:
        tree.batch_search(dq, nq, counts, offs, out, total)
        torch.cuda.synchronize()
    return counts.cpu().numpy(), offs.cpu().numpy(), out


def _check_batch(P, pts, qs, want, c, o, out, sample):
    assert np.array_equal(c, want), np.flatnonzero(c != want)[:8]
    assert o[len(qs)] == int(want.sum())
    for k in sample:
        got = np.sort(out[o[k]:o[k] + c[k]].cpu().numpy().view(np.uint32))
        assert np.array_equal(got, _brute(pts, qs[k])), k


def test_2e26_float_tree_and_mixed_batch(torch, P):
    import pst_data as D
    n = 1 << 26
    pts = D.interval_points(n, seed=0)
    d = P.to_device(pts)
    tree = P.StaticPSTGPU(d, n, "f32", "u32")
    del d
    torch.cuda.empty_cache()
    print("2^26 tree built; whole-tree invariants", flush=True)
    TI.check(tree.export_tree(), pts)
    # 256 two-sided (half left, half right (q, q)) and 256 three-sided queries
    qs = D.float_query_mix(256, 256, seed=0)
    kinds = qs["kind"].astype(np.int32)
    tr = _oracle_tree(P, tree, "f32", n)
    want = O.count_queries(tr, kinds, qs["min_dim1"], qs["max_dim1"], qs["min_dim2"], threads=THREADS)
    del tr
    total = int(want.sum())
    print(f"oracle counts done: {total} IDs, {int(np.sum(kinds == 2))} right queries", flush=True)
    right = np.flatnonzero(kinds == 2)[:4]
    left = np.flatnonzero(kinds == 1)[:4]
    three = np.flatnonzero(kinds == 0)[:8]
    sample = np.concatenate([right, left, three])
    c, o, out = _run_batch(torch, P, tree, qs, total)
    _check_batch(P, pts, qs, want, c, o, out, sample)
    print("walk batch equal; level index", flush=True)
    del out
    torch.cuda.empty_cache()
    tree.set_level_index(True)
    c, o, out = _run_batch(torch, P, tree, qs, total)
    _check_batch(P, pts, qs, want, c, o, out, sample)
    del out
    tree.close()
    torch.cuda.empty_cache()


def test_2e27_int32_graph_batch(torch, P):
    import pst_data as D
    n = 1 << 27
    pts = D.int_points(n, seed=3)
    d = P.to_device(pts)
    tree = P.StaticPSTGPU(d, n, "i32", "u32")
    del d
    torch.cuda.empty_cache()
    qs = D.int_query_mix(8192, seed=3)
    kinds = qs["kind"].astype(np.int32)
    tr = _oracle_tree(P, tree, "i32", n)
    want = O.count_queries(tr, kinds, qs["min_dim1"], qs["max_dim1"], qs["min_dim2"], threads=THREADS)
    del tr
    total = int(want.sum())
    print(f"2^27 int32 tree built, oracle counts done: {total} IDs", flush=True)
    rng = np.random.default_rng(5)
    sample = np.sort(rng.choice(len(qs), 16, replace=False))
    cfg = tree.autotune(P.to_device(qs), len(qs))
    print(f"autotuned {cfg}", flush=True)
    c, o, out = _run_batch(torch, P, tree, qs, total, graph=True)
    _check_batch(P, pts, qs, want, c, o, out, sample)
    del out
    print("walk graph equal; level index", flush=True)
    tree.set_level_index(True)
    c, o, out = _run_batch(torch, P, tree, qs, total, graph=True)
    _check_batch(P, pts, qs, want, c, o, out, sample)
    del out
    tree.close()
    torch.cuda.empty_cache()

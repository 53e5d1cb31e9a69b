"""pst_gpu_batch_search_noalloc, called through the C ABI: the same counts,
offsets and output bytes as pst_gpu_batch_search through the level index,
on two calls in a row (the second smaller, so stale scratch contents of the
first would show), a sizing call that writes nothing, and the documented
refusals."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

GUARD = 0x3C3C3C3C


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


@pytest.fixture(scope="module")
def P():
    import pst_amd
    L = pst_amd.lib()
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    L.pst_gpu_batch_search_noalloc.argtypes = [vp, vp, sz, vp, vp, vp, sz, vp]
    return pst_amd


def _noalloc(P, tr, dq, nq, counts, offs, out, cap, stream=None):
    return P.lib().pst_gpu_batch_search_noalloc(tr.h, dq.data_ptr(), nq, counts.data_ptr(), offs.data_ptr(),
                                                out.data_ptr() if out is not None else None, cap,
                                                P._stream_ptr(stream))


def _run_pair(torch, P, tr, qs):
    nq = len(qs)
    dq = P.to_device(qs)
    c1 = torch.zeros(nq, dtype=torch.int64, device="cuda")
    o1 = torch.zeros(nq + 1, dtype=torch.int64, device="cuda")
    tr.batch_count(dq, nq, c1)
    torch.cuda.synchronize()
    total = int(c1.sum().item())
    out1 = torch.full((total + 64,), GUARD, dtype=torch.int32, device="cuda")
    tr.batch_search(dq, nq, c1, o1, out1, total)
    c2 = torch.full((nq,), -1, dtype=torch.int64, device="cuda")
    o2 = torch.full((nq + 1,), -1, dtype=torch.int64, device="cuda")
    out2 = torch.full((total + 64,), GUARD, dtype=torch.int32, device="cuda")
    rc = _noalloc(P, tr, dq, nq, c2, o2, out2, total)
    assert rc == 0, P.lib().pst_last_error()
    torch.cuda.synchronize()
    assert torch.equal(c1, c2) and torch.equal(o1, o2)
    assert torch.equal(out1, out2), "output bytes differ from pst_gpu_batch_search"
    return c2.cpu().numpy(), o2.cpu().numpy(), out2.cpu().numpy().view(np.uint32)


def test_noalloc_equals_batch_search_twice(torch, P):
    import pst_data as D
    n = (1 << 20) + 77
    pts = D.interval_points(n, seed=51)
    tr = P.StaticPSTGPU(P.to_device(pts), n, "f32", "u32")
    tr.set_level_index(True)
    for call, (n3, n2, seed) in enumerate(((1024, 1024, 52), (100, 60, 53))):
        qs = D.float_query_mix(n3, n2, seed=seed)
        print(f"noalloc call {call}: {len(qs)} queries", flush=True)
        c, o, r = _run_pair(torch, P, tr, qs)
        for k in range(0, len(qs), max(1, len(qs) // 12)):
            q = qs[k]
            exp = pts["id"][O.brute_force(pts, int(q["kind"]), q["min_dim1"], q["max_dim1"], q["min_dim2"])]
            assert np.array_equal(np.sort(r[o[k]:o[k] + c[k]]), np.sort(exp)), (call, k)
    tr.close()


def test_noalloc_int32_tree_and_sizing_call(torch, P):
    import pst_data as D
    n = 300000
    pts = D.int_points(n, seed=61)
    tr = P.StaticPSTGPU(P.to_device(pts), n, "i32", "u32")
    tr.set_level_index(True)
    qs = D.int_query_mix(8192, seed=62)   # the most queries one call holds
    print("noalloc int32 8192 queries", flush=True)
    _run_pair(torch, P, tr, qs)
    nq = len(qs)
    dq = P.to_device(qs)
    counts = torch.zeros(nq, dtype=torch.int64, device="cuda")
    offs = torch.zeros(nq + 1, dtype=torch.int64, device="cuda")
    out = torch.full((1024,), GUARD, dtype=torch.int32, device="cuda")
    assert _noalloc(P, tr, dq, nq, counts, offs, out, 0) == 0
    torch.cuda.synchronize()
    assert bool((out == GUARD).all().item())
    assert int(offs[nq].item()) == int(counts.sum().item())
    # a truncating capacity writes exactly the first `cap` records of the full output
    cap = 777
    ref_c = torch.zeros_like(counts)
    ref_o = torch.zeros_like(offs)
    ref = torch.full((1024,), GUARD, dtype=torch.int32, device="cuda")
    tr.batch_search(dq, nq, ref_c, ref_o, ref, cap)
    assert _noalloc(P, tr, dq, nq, counts, offs, out, cap) == 0
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    tr.close()


def test_noalloc_refusals(torch, P):
    import pst_data as D
    n = 5000
    pts = D.interval_points(n, seed=71)
    tr = P.StaticPSTGPU(P.to_device(pts), n, "f32", "u32")
    qs = D.float_query_mix(8, 8, seed=72)
    dq = P.to_device(qs)
    c = torch.zeros(8193, dtype=torch.int64, device="cuda")
    o = torch.zeros(8194, dtype=torch.int64, device="cuda")
    out = torch.zeros(16, dtype=torch.int32, device="cuda")
    assert _noalloc(P, tr, dq, 16, c, o, out, 16) == 1          # no level index yet
    assert b"level index" in P.lib().pst_last_error()
    tr.set_level_index(True)
    big = P.to_device(D.float_query_mix(4097, 4096, seed=73))
    assert _noalloc(P, tr, big, 8193, c, o, out, 16) == 1      # beyond the scratch's 8192 queries
    assert b"8192" in P.lib().pst_last_error()
    tr.close()
    f = P.StaticPSTGPUArr(P.to_device(pts), n, 64, "f32", "u32")
    assert _noalloc(P, f, dq, 16, c, o, out, 16) == 1           # forests are not served
    f.close()

"""Tree persistence and the raw volume reader on the GPU: a saved and reloaded
tree or forest exports bit for bit the same records as the original (and so
the oracle's tree), answers searches and batches identically, and a file
with a changed byte fails its checksum. Raw volume files reach device
memory byte for byte and feed metacell formation."""
import os

import numpy as np
import pytest

import golden_util as G
import oracle as O

pytestmark = pytest.mark.gpu


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


def _points(rng, t, i, n):
    p = np.zeros(n, O.point_dtype(t, i))
    if t in ("f32", "f64"):
        p["dim1"] = rng.random(n)
        p["dim2"] = p["dim1"] + rng.exponential(0.05, n)
    else:
        p["dim1"] = rng.integers(-10 ** 6, 10 ** 6, n)
        p["dim2"] = p["dim1"] + rng.integers(0, 10 ** 5, n)
    if i is not None:
        p["id"] = rng.permutation(n)
    return p


def _same_export(a, b):
    ea, eb = a.export_tree(), b.export_tree()
    assert set(ea) == set(eb)
    for k in ea:
        assert np.array_equal(G.bits_arr(ea[k]), G.bits_arr(eb[k])), k


@pytest.mark.parametrize("t,i,n", [("f32", "u32", 40000), ("f64", None, 5000), ("i32", "u64", 4099),
                                   ("i64", "i64", 1), ("f32", "u32", 0)])
def test_tree_round_trip(torch, P, tmp_path, t, i, n):
    rng = np.random.default_rng(n + 3)
    pts = _points(rng, t, i, n)
    print(f"persist tree {t}/{i} n={n}", flush=True)
    d = P.to_device(pts) if n else None
    tr = P.StaticPSTGPU(d.data_ptr() if n else 0, n, t, i)
    path = str(tmp_path / "tree.pst")
    tr.save(path)
    tr.save(str(tmp_path / "again.pst"))
    with open(path, "rb") as f1, open(tmp_path / "again.pst", "rb") as f2:
        assert f1.read() == f2.read()  # a file depends only on the tree
    lo = P.StaticPSTGPU.load(path)
    assert type(lo) is P.StaticPSTGPU and (lo.t, lo.i, lo.n) == (t, i, n)
    if n == 0:
        return
    _same_export(tr, lo)
    ora = O.build(pts, t, i)
    for k, v in ora.arrays().items():
        assert np.array_equal(G.bits_arr(lo.export_tree()[k]), G.bits_arr(v)), k
    lo1, hi1 = float(pts["dim1"].min()), float(pts["dim2"].max())
    for q in np.linspace(lo1, hi1, 7):
        for kind in (O.THREE, O.LEFT, O.RIGHT):
            a, b, c = q, q + (hi1 - lo1) / 5, q
            rp = P.RET_POINT if i is None else P.RET_ID
            f = {O.THREE: "threeSidedSearch", O.LEFT: "twoSidedLeftSearch", O.RIGHT: "twoSidedRightSearch"}[kind]
            args = (a, b, c) if kind == O.THREE else ((b, c) if kind == O.LEFT else (a, c))
            want = getattr(tr, f)(*args, ret=rp)
            got = getattr(lo, f)(*args, ret=rp)
            if rp == P.RET_ID:
                assert np.array_equal(np.sort(got), np.sort(want))
            else:
                assert G.canon(got, True, i is not None) == G.canon(want, True, i is not None)
    lo.close()
    tr.close()


@pytest.mark.parametrize("tpb,n", [(64, 10000), (1000, 3001)])
def test_forest_round_trip(torch, P, tmp_path, tpb, n):
    rng = np.random.default_rng(tpb + n)
    pts = _points(rng, "f32", "u32", n)
    print(f"persist forest tpb={tpb} n={n}", flush=True)
    d = P.to_device(pts)
    f = P.StaticPSTGPUArr(d.data_ptr(), n, tpb, "f32", "u32")
    path = str(tmp_path / "forest.pst")
    f.save(path)
    lo = P.StaticPSTGPU.load(path)
    assert type(lo) is P.StaticPSTGPUArr and lo.slots_per_tree == f.slots_per_tree
    assert lo.forest_sizes() == f.forest_sizes()
    _same_export(f, lo)
    for q in rng.random(8).astype(np.float32):
        want = np.sort(f.twoSidedLeftSearch(q, q))
        got = np.sort(lo.twoSidedLeftSearch(q, q))
        assert np.array_equal(got, want)
        assert np.array_equal(got, np.sort(pts["id"][(pts["dim1"] <= q) & (q <= pts["dim2"])]))
    lo.close()
    f.close()


def test_loaded_tree_batches_like_original(torch, P, tmp_path):
    """Batched count + write on a reloaded tree (its side arrays rebuilt on
    load) equals the original's, with whole-subtree copies on."""
    import pst_data as D
    n = 1 << 18
    pts = D.interval_points(n, seed=11)
    d = P.to_device(pts)
    tr = P.StaticPSTGPU(d.data_ptr(), n, "f32", "u32")
    path = str(tmp_path / "t.pst")
    tr.save(path)
    lo = P.StaticPSTGPU.load(path)
    qs = D.float_query_mix(64, 64, seed=12)
    nq = len(qs)
    qd = P.to_device(qs)
    res = []
    for t in (tr, lo):
        t.set_subtree_copy(16, 65535)
        counts = torch.zeros(nq, dtype=torch.int64, device="cuda")
        offs = torch.zeros(nq + 1, dtype=torch.int64, device="cuda")
        out = torch.zeros(n * 64, dtype=torch.int32, device="cuda")
        t.batch_search(qd, nq, counts, offs, out, n * 64)
        torch.cuda.synchronize()
        c, o, r = counts.cpu().numpy(), offs.cpu().numpy(), out.cpu().numpy()
        res.append([np.sort(r[o[k]:o[k] + c[k]]) for k in range(nq)])
    for k in range(nq):
        assert np.array_equal(res[0][k], res[1][k]), k
    lo.close()
    tr.close()


def test_changed_byte_fails_checksum(torch, P, tmp_path):
    rng = np.random.default_rng(5)
    pts = _points(rng, "f32", "u32", 3000)
    d = P.to_device(pts)
    tr = P.StaticPSTGPU(d.data_ptr(), len(pts), "f32", "u32")
    path = str(tmp_path / "t.pst")
    tr.save(path)
    raw = bytearray(open(path, "rb").read())
    raw[64 + 4 * 1000] ^= 0x01  # one bit of dim1 of slot 1000
    open(path, "wb").write(bytes(raw))
    with pytest.raises(P.PSTError) as e:
        P.StaticPSTGPU.load(path)
    assert e.value.code == P.ERR_IO and "checksum" in str(e.value)
    tr.close()


@pytest.mark.parametrize("dims,t", [((128, 128, 128), "f32"), ((37, 29, 41), "f64"), ((50, 6, 13), "i32")])
def test_raw_volume_to_device_and_metacells(torch, P, tmp_path, dims, t):
    """readInVertices semantics: the device copy equals the file; metacells
    formed from it equal those of the same volume copied directly."""
    rng = np.random.default_rng(sum(dims))
    px, py, pz = dims
    vol = (rng.standard_normal(px * py * pz) * 100).astype(O.NP[t])
    path = str(tmp_path / "vol.raw")
    vol.tofile(path)
    print(f"raw volume {dims} {t}", flush=True)
    dv = torch.zeros(vol.nbytes, dtype=torch.uint8, device="cuda")
    P.read_vertices_to_device(path, vol.size, dv, dtype=t)
    torch.cuda.synchronize()
    assert np.array_equal(dv.cpu().numpy().view(O.NP[t]), vol)
    n, g = P.calcNumMetacells(dims, (4, 4, 4))
    rec = P.point_dtype(t, "u32")
    a = torch.zeros(n * rec.itemsize, dtype=torch.uint8, device="cuda")
    b = torch.zeros_like(a)
    P.form_metacells_into(dv, dims, (4, 4, 4), a, dtype=t, id_dtype="u32")
    P.form_metacells_into(P.to_device(vol), dims, (4, 4, 4), b, dtype=t, id_dtype="u32")
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    os.remove(path)

"""Marching cubes over active metacells on the GPU (pst_marching_cubes_*):

* against the Python restatement of the case rule (tests/mc_table.py):
  the same triangles, bit for bit and in order, per metacell, for three
  types and four metacell shapes with partial ID lists;
* against geometry: a sphere's distance field gives the sphere's area and
  radius with outward normals and a closed, consistently wound mesh of
  Euler characteristic 2; a torus gives Euler characteristic 0;
* end to end: volume -> metacells -> forest search (q, q) -> IDs -> mesh
  equals the mesh of all metacells.
"""
import numpy as np
import pytest

import mc_table as M
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


def ref_metacell(vol, mc, cell, q):
    pz, py, px = vol.shape
    mx, my, mz = mc
    gx = -(-(px - 1) // mx)
    gy = -(-(py - 1) // my)
    a, b, c = cell % gx, (cell // gx) % gy, cell // (gx * gy)
    acc = np.float32 if vol.dtype == np.float32 else np.float64
    out = []
    for z in range(c * mz, min((c + 1) * mz, pz - 1)):
        for y in range(b * my, min((b + 1) * my, py - 1)):
            for x in range(a * mx, min((a + 1) * mx, px - 1)):
                f = [vol[z + (k >> 2), y + ((k >> 1) & 1), x + (k & 1)] for k in range(8)]
                out.extend(M.voxel_triangles(f, q, x, y, z, acc))
    return out


def _mesh(P, torch, vol, dims, mc, ids, q, dtype="f32"):
    return P.isosurface(torch.from_numpy(vol.ravel()).cuda(), dims, mc,
                        torch.from_numpy(np.asarray(ids, np.uint32).view(np.int32)).cuda(), len(ids), q,
                        dtype=dtype, method="cubes")


@pytest.mark.parametrize("dims,mc,t", [((13, 11, 9), (4, 4, 4), "f32"), ((10, 7, 12), (3, 2, 5), "f64"),
                                       ((9, 9, 9), (1, 1, 1), "i32"), ((6, 5, 4), (8, 8, 8), "f32")])
def test_equals_python_restatement(torch, P, dims, mc, t):
    rng = np.random.default_rng(sum(dims) * 3 + sum(mc))
    px, py, pz = dims
    if t == "i32":
        vol = rng.integers(-50, 50, (pz, py, px)).astype(np.int32)
        q = 3
    else:
        vol = rng.standard_normal((pz, py, px)).astype(O.NP[t])
        q = 0.1
    n, g = P.calcNumMetacells(dims, mc)
    ids = np.array([k for k in range(n) if k % 3 != 1], np.uint32)
    print(f"marching cubes {dims} {mc} {t}", flush=True)
    vd = torch.from_numpy(vol.ravel()).cuda()
    idd = torch.from_numpy(ids.view(np.int32)).cuda()
    counts = torch.zeros(len(ids), dtype=torch.int64, device="cuda")
    P.isosurface_count(vd, dims, mc, idd, len(ids), q, counts, dtype=t, method="cubes")
    torch.cuda.synchronize()
    cnt = counts.cpu().numpy()
    offs = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    tri = torch.zeros(max(1, int(offs[-1])) * 9, dtype=torch.float32, device="cuda")
    P.isosurface_write(vd, dims, mc, idd, len(ids), q, torch.from_numpy(offs).cuda(), tri, int(offs[-1]), dtype=t,
                       method="cubes")
    torch.cuda.synchronize()
    got = tri.cpu().numpy()[:int(offs[-1]) * 9].reshape(-1, 3, 3)
    for k, cell in enumerate(ids):
        want = ref_metacell(vol, mc, int(cell), q)
        assert cnt[k] == len(want), cell
        mine = got[offs[k]:offs[k + 1]]
        assert all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(mine, want)), cell
    blocking = _mesh(P, torch, vol, dims, mc, ids, q, dtype=t)
    assert np.array_equal(blocking, got)


def _closed_euler(tri):
    verts = {}
    idx = np.array([[verts.setdefault(tuple(v), len(verts)) for v in t3] for t3 in tri.tolist()])
    directed = {}
    for a, b, c in idx:
        for e in ((a, b), (b, c), (c, a)):
            directed[e] = directed.get(e, 0) + 1
    assert all(v == 1 for v in directed.values()), "an edge is used twice in one direction"
    assert all((b, a) in directed for (a, b) in directed), "the mesh has a boundary"
    V, E, F = len(verts), len(directed) // 2, len(idx)
    return V - E + F


def test_sphere_area_radius_normals_closed(torch, P):
    side, c, r = 48, np.array([23.37, 23.61, 23.12]), 15.3
    k, j, i = np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij")
    vol = np.sqrt((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2).astype(np.float32)
    dims, mc = (side, side, side), (4, 4, 4)
    n, _ = P.calcNumMetacells(dims, mc)
    tri = _mesh(P, torch, vol, dims, mc, np.arange(n), r)
    assert len(tri) > 1000
    tri64 = tri.astype(np.float64)
    nrm = np.cross(tri64[:, 1] - tri64[:, 0], tri64[:, 2] - tri64[:, 0])
    area = 0.5 * np.linalg.norm(nrm, axis=1).sum()
    assert abs(area / (4 * np.pi * r * r) - 1) < 0.01, area
    dist = np.linalg.norm(tri64.reshape(-1, 3) - c, axis=1)
    assert np.max(np.abs(dist - r)) < 0.05
    cen = tri64.mean(axis=1) - c
    big = np.linalg.norm(nrm, axis=1) > 1e-9
    assert np.all(np.einsum("ij,ij->i", nrm[big], cen[big]) > 0)  # outward: toward larger distance
    assert _closed_euler(tri) == 2


def test_torus_is_closed_with_euler_characteristic_zero(torch, P):
    side, R, r = 56, 16.0, 6.5
    c = np.array([27.3, 27.6, 27.1])
    k, j, i = np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij")
    x, y, z = i - c[0], j - c[1], k - c[2]
    vol = (np.sqrt((np.sqrt(x ** 2 + y ** 2) - R) ** 2 + z ** 2)).astype(np.float32)
    dims, mc = (side, side, side), (4, 4, 4)
    n, _ = P.calcNumMetacells(dims, mc)
    tri = _mesh(P, torch, vol, dims, mc, np.arange(n), r)
    assert _closed_euler(tri) == 0
    tri64 = tri.astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(tri64[:, 1] - tri64[:, 0], tri64[:, 2] - tri64[:, 0]), axis=1).sum()
    assert abs(area / (4 * np.pi ** 2 * R * r) - 1) < 0.01, area


def test_search_ids_give_the_whole_surface(torch, P):
    dims, mc = (65, 65, 65), (4, 4, 4)
    prm = P.gaussian_volume_params(8, 2)
    vol = torch.empty(65 ** 3, dtype=torch.float32, device="cuda")
    P.synth_gaussian_volume(vol, dims, torch.from_numpy(prm.ravel()).cuda(), 8, 0.01, 2)
    n, g = P.calcNumMetacells(dims, mc)
    rec = P.point_dtype("f32", "u32")
    cells = torch.zeros(n * rec.itemsize, dtype=torch.uint8, device="cuda")
    P.form_metacells_into(vol, dims, mc, cells)
    torch.cuda.synchronize()
    f = P.StaticPSTGPUArr(cells.data_ptr(), n, 64, "f32", "u32")
    allids = torch.arange(n, dtype=torch.int32, device="cuda")
    h = cells.cpu().numpy().view(rec)
    key = lambda tr: tuple(sorted(tuple(v.tolist()) for v in tr))  # noqa: E731
    for q in np.linspace(float(h["dim1"].min()), float(h["dim2"].max()), 5)[1:-1].astype(np.float32):
        ids = np.sort(f.twoSidedLeftSearch(q, q))
        part = P.isosurface(vol, dims, mc, torch.from_numpy(ids.view(np.int32)).cuda(), len(ids), q, method="cubes")
        whole = P.isosurface(vol, dims, mc, allids, n, q, method="cubes")
        assert len(part) == len(whole) > 0
        assert sorted(key(x) for x in part) == sorted(key(x) for x in whole)
    f.close()

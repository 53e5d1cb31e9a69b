"""Isosurface extraction over active metacells (marching tetrahedra) on the
GPU, checked three ways:

* against a numpy restatement of the same scheme (six tetrahedra per voxel,
  vertices interpolated from the lower-indexed edge endpoint, winding toward
  values >= q): the same triangles, bit for bit, per metacell;
* against geometry: the isosurface of a sphere's distance field has the
  sphere's area and radius, outward normals, and is closed (every edge in
  exactly two triangles, once in each direction);
* end to end: volume -> metacells -> forest search (q, q) -> IDs -> mesh
  equals the mesh of all metacells.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TETS = [(0, 1, 3, 7), (0, 1, 5, 7), (0, 2, 3, 7), (0, 2, 6, 7), (0, 4, 5, 7), (0, 4, 6, 7)]


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
    """Triangles of one metacell, in the order the kernel writes them: chunks
    of 64 voxels (voxels x fastest), within a chunk tetrahedron by
    tetrahedron (TETS order), within that voxel by voxel."""
    pz, py, px = vol.shape
    mx, my, mz = mc
    gx = -(-(px - 1) // mx)
    gy = -(-(py - 1) // my)
    a, b, c = cell % gx, (cell // gx) % gy, cell // (gx * gy)
    acc = np.float32 if vol.dtype == np.float32 else np.float64
    qa = acc(q)
    voxels = [(x, y, z) for z in range(c * mz, min((c + 1) * mz, pz - 1))
              for y in range(b * my, min((b + 1) * my, py - 1))
              for x in range(a * mx, min((a + 1) * mx, px - 1))]
    out = []
    for j0 in range(0, len(voxels), 64):
        for tet in TETS:
            for (x, y, z) in voxels[j0:j0 + 64]:
                f = [vol[z + (k >> 2), y + ((k >> 1) & 1), x + (k & 1)] for k in range(8)]
                pos = [np.array([x + (k & 1), y + ((k >> 1) & 1), z + (k >> 2)], np.float32) for k in range(8)]

                def edge(i, j):
                    u, w = min(i, j), max(i, j)
                    t = (qa - acc(f[u])) / (acc(f[w]) - acc(f[u]))
                    return pos[u] + np.float32(t) * (pos[w] - pos[u])

                below = [k for k in tet if f[k] < q]
                above = [k for k in tet if not f[k] < q]
                if not below or not above:
                    continue
                up = pos[above[0]] - pos[below[0]]
                if len(below) == 1 or len(above) == 1:
                    s = below[0] if len(below) == 1 else above[0]
                    o = above if len(below) == 1 else below
                    tris = [(edge(s, o[0]), edge(s, o[1]), edge(s, o[2]))]
                else:
                    p0, p1 = edge(below[0], above[0]), edge(below[0], above[1])
                    p2, p3 = edge(below[1], above[1]), edge(below[1], above[0])
                    tris = [(p0, p1, p2), (p0, p2, p3)]
                for ta, tb, tc in tris:
                    n = np.cross(tb - ta, tc - ta)
                    if float(np.dot(n, up)) < 0:
                        tb, tc = tc, tb
                    out.append(np.array((ta, tb, tc), np.float32))
    return out


def tri_key(tr):
    """A triangle as its sorted vertices (winding compared separately)."""
    return tuple(sorted(tuple(v.tolist()) for v in tr))


@pytest.mark.parametrize("dims,mc,t", [((13, 11, 9), (4, 4, 4), "f32"), ((10, 7, 12), (3, 2, 5), "f64"),
                                       ((9, 9, 9), (1, 1, 1), "i32"), ((6, 5, 4), (8, 8, 8), "f32")])
def test_equals_numpy_restatement(torch, P, dims, mc, t):
    rng = np.random.default_rng(sum(dims) + sum(mc))
    px, py, pz = dims
    if t == "i32":
        vol = rng.integers(-50, 50, (pz, py, px)).astype(np.int32)
        q = 3
    else:
        vol = rng.standard_normal((pz, py, px)).astype(O.NP[t])
        q = 0.1
    n, g = P.calcNumMetacells(dims, mc)
    ids = np.array([k for k in range(n) if k % 3 != 1], np.uint32)  # a partial list
    print(f"isosurface numpy {dims} {mc} {t}", flush=True)
    vd = torch.from_numpy(vol.ravel()).cuda()
    idd = torch.from_numpy(ids.view(np.int32)).cuda()
    counts = torch.zeros(len(ids), dtype=torch.int64, device="cuda")
    P.isosurface_count(vd, dims, mc, idd, len(ids), q, counts, dtype=t)
    torch.cuda.synchronize()
    cnt = counts.cpu().numpy()
    offs = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    tri = torch.zeros(max(1, int(offs[-1])) * 9, dtype=torch.float32, device="cuda")
    P.isosurface_write(vd, dims, mc, idd, len(ids), q, torch.from_numpy(offs).cuda(), tri, int(offs[-1]), dtype=t)
    torch.cuda.synchronize()
    got = tri.cpu().numpy()[:int(offs[-1]) * 9].reshape(-1, 3, 3)
    for k, cell in enumerate(ids):
        want = ref_metacell(vol, mc, int(cell), q)
        assert cnt[k] == len(want), cell
        mine = got[offs[k]:offs[k + 1]]
        assert [tri_key(x) for x in mine] == [tri_key(x) for x in want], cell
        # winding: same as the restatement wherever the triangle is not degenerate
        for a, b in zip(mine, want):
            n = np.cross(b[1] - b[0], b[2] - b[0])
            if np.linalg.norm(n) > 1e-6:
                assert np.array_equal(a, b), cell
    blocking = P.isosurface(vd, dims, mc, idd, len(ids), q, dtype=t)
    assert np.array_equal(blocking, got)


def test_sphere_area_radius_normals_closed(torch, P):
    side, c, r = 48, np.array([23.37, 23.61, 23.12]), 15.3
    k, j, i = np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij")
    vol = np.sqrt((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2).astype(np.float32)
    dims, mc = (side, side, side), (4, 4, 4)
    n, _ = P.calcNumMetacells(dims, mc)
    ids = np.arange(n, dtype=np.uint32)
    tri = P.isosurface(torch.from_numpy(vol.ravel()).cuda(), dims, mc, torch.from_numpy(ids.view(np.int32)).cuda(),
                       n, r)
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
    # closed and consistently wound: every directed edge once, its reverse once
    verts = {}
    idx = np.array([[verts.setdefault(tuple(v), len(verts)) for v in t3] for t3 in tri.tolist()])
    directed = {}
    for a, b, cc in idx:
        for e in ((a, b), (b, cc), (cc, a)):
            directed[e] = directed.get(e, 0) + 1
    assert all(v == 1 for v in directed.values())
    assert all((b, a) in directed for (a, b) in directed)


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
    for q in np.linspace(float(h["dim1"].min()), float(h["dim2"].max()), 5)[1:-1].astype(np.float32):
        ids = np.sort(f.twoSidedLeftSearch(q, q))
        part = P.isosurface(vol, dims, mc, torch.from_numpy(ids.view(np.int32)).cuda(), len(ids), q)
        whole = P.isosurface(vol, dims, mc, allids, n, q)
        assert len(part) == len(whole) > 0
        assert sorted(tri_key(x) for x in part) == sorted(tri_key(x) for x in whole)
    f.close()


def test_id_past_grid_rejected(torch, P):
    dims, mc = (9, 9, 9), (4, 4, 4)
    vol = torch.zeros(9 ** 3, dtype=torch.float32, device="cuda")
    bad = torch.tensor([0, 8], dtype=torch.int32, device="cuda")  # 2 x 2 x 2 metacells: 8 is past the grid
    with pytest.raises(P.PSTError) as e:
        P.isosurface(vol, dims, mc, bad, 2, 0.5)
    assert e.value.code == 1 and "past the metacell grid" in str(e.value)

"""GPU metacell formation (formMetacellTags) against a brute-force min/max
over each metacell's vertex box, the seeded synthetic volume against a
float64 evaluation of its formula, and an end-to-end isovalue query on a
forest built from metacells."""
import numpy as np
import pytest

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


def brute_metacells(vol, mc):
    """vol[z, y, x]; metacell (a, b, c) covers vertices [a*mx, min((a+1)*mx, px-1)]."""
    pz, py, px = vol.shape
    mx, my, mz = mc
    gx, gy, gz = [-(-(p - 1) // m) for p, m in ((px, mx), (py, my), (pz, mz))]
    mn = np.empty((gz, gy, gx), vol.dtype)
    mxv = np.empty((gz, gy, gx), vol.dtype)
    for c in range(gz):
        for b in range(gy):
            for a in range(gx):
                box = vol[c * mz:min((c + 1) * mz, pz - 1) + 1, b * my:min((b + 1) * my, py - 1) + 1,
                          a * mx:min((a + 1) * mx, px - 1) + 1]
                mn[c, b, a] = box.min()
                mxv[c, b, a] = box.max()
    return mn.ravel(), mxv.ravel(), (gx, gy, gz)


@pytest.mark.parametrize("dims,mc,t", [((37, 29, 41), (4, 4, 4), "f32"), ((20, 11, 9), (3, 5, 2), "f32"),
                                       ((9, 8, 7), (1, 1, 1), "f32"), ((50, 6, 13), (7, 3, 9), "f64"),
                                       ((33, 17, 5), (4, 4, 4), "i32"), ((2, 2, 2), (4, 4, 4), "f32")])
def test_metacells_equal_brute_force(torch, P, dims, mc, t):
    rng = np.random.default_rng(sum(dims) + sum(mc))
    px, py, pz = dims
    if t == "i32":
        vol = rng.integers(-1000, 1000, (pz, py, px)).astype(np.int32)
    else:
        vol = rng.standard_normal((pz, py, px)).astype(O.NP[t])
    n, g = P.calcNumMetacells(dims, mc)
    mn, mxv, g2 = brute_metacells(vol, mc)
    assert tuple(g) == g2 and n == len(mn)
    vd = torch.from_numpy(vol.ravel()).cuda()
    rec = P.point_dtype(t, "u32")
    out = torch.zeros(n * rec.itemsize, dtype=torch.uint8, device="cuda")
    P.form_metacells_into(vd, dims, mc, out, dtype=t, id_dtype="u32")
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(rec)
    assert np.array_equal(got["dim1"], mn)
    assert np.array_equal(got["dim2"], mxv)
    assert np.array_equal(got["id"], np.arange(n, dtype=np.uint32))


def splitmix64(x):
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def test_synthetic_volume_formula(torch, P):
    """Float kernel: within 2e-4 absolute of a float64 evaluation of the
    same formula (f32 exp/log/cos and a 16-term f32 sum)."""
    dims = (23, 17, 11)
    ng, seed, std = 16, 7, 0.01
    prm = P.gaussian_volume_params(ng, seed)
    pd = torch.from_numpy(prm.ravel()).cuda()
    vol = torch.empty(dims[0] * dims[1] * dims[2], dtype=torch.float32, device="cuda")
    P.synth_gaussian_volume(vol, dims, pd, ng, std, seed)
    torch.cuda.synchronize()
    got = vol.cpu().numpy().astype(np.float64)
    px, py, pz = dims
    k, j, i = np.meshgrid(np.arange(pz), np.arange(py), np.arange(px), indexing="ij")
    x = (i.astype(np.float32) * np.float32(1.0 / (px - 1))).astype(np.float64).ravel()
    y = (j.astype(np.float32) * np.float32(1.0 / (py - 1))).astype(np.float64).ravel()
    z = (k.astype(np.float32) * np.float32(1.0 / (pz - 1))).astype(np.float64).ravel()
    want = np.zeros_like(x)
    for gi in range(ng):
        cx, cy, cz, w, amp = prm[gi, :5].astype(np.float64)
        want += amp * np.exp(((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) * w)
    v = np.arange(len(x), dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * np.uint64(0x100000001B3)
        h1 = splitmix64(base + np.uint64(2) * v)
        h2 = splitmix64(base + np.uint64(2) * v + np.uint64(1))
    u1 = ((h1 >> np.uint64(40)).astype(np.float64) + 1) / 16777217.0
    u2 = (h2 >> np.uint64(40)).astype(np.float64) / 16777216.0
    want += std * np.sqrt(-2 * np.log(u1)) * np.cos(2 * np.pi * u2)
    assert np.max(np.abs(got - want)) < 2e-4


def test_isovalue_queries_on_metacell_forest(torch, P):
    """Volume -> metacells -> forest (threads_per_block 64) -> isovalue
    queries (q, q): the reported metacell IDs are those whose [min, max]
    contains q."""
    dims = (65, 65, 65)
    prm = P.gaussian_volume_params(8, 1)
    pd = torch.from_numpy(prm.ravel()).cuda()
    vol = torch.empty(65 ** 3, dtype=torch.float32, device="cuda")
    P.synth_gaussian_volume(vol, dims, pd, 8, 0.01, 1)
    n, g = P.calcNumMetacells(dims, (4, 4, 4))
    rec = P.point_dtype("f32", "u32")
    mc = torch.zeros(n * rec.itemsize, dtype=torch.uint8, device="cuda")
    P.form_metacells_into(vol, dims, (4, 4, 4), mc)
    torch.cuda.synchronize()
    cells = mc.cpu().numpy().view(rec)
    f = P.StaticPSTGPUArr(mc.data_ptr(), n, 64, "f32", "u32")
    lo, hi = float(cells["dim1"].min()), float(cells["dim2"].max())
    for q in np.linspace(lo, hi, 9).astype(np.float32):
        got = np.sort(f.twoSidedLeftSearch(q, q))
        want = np.sort(cells["id"][(cells["dim1"] <= q) & (q <= cells["dim2"])])
        assert np.array_equal(got, want), q
    f.close()

"""The marching cubes case rule, checked on the CPU for all 256 corner
configurations: triangles use exactly the edges whose endpoints differ, the
patch of every configuration is a surface whose boundary lies on the cube's
faces with each face crossed by segments between that face's crossing edges,
normals point from below to above, and no configuration needs more than 5
triangles."""
import itertools

import numpy as np

import mc_table as M


def test_edges_used_are_the_crossing_edges():
    for cfg in range(256):
        cross = {e for e, (a, b) in enumerate(M.EDGES) if ((cfg >> a) & 1) != ((cfg >> b) & 1)}
        used = {e for t in M.TABLE[cfg] for e in t}
        assert used == cross, cfg


def test_at_most_five_triangles():
    assert max(len(t) for t in M.TABLE) == 5
    assert len(M.TABLE[0]) == len(M.TABLE[255]) == 0


def test_patch_boundary_lies_on_faces_and_interior_edges_pair_up():
    """Interior edges of a patch (between two of its triangles) appear once in
    each direction; boundary edges join two crossing edges of one face."""
    for cfg in range(256):
        directed = {}
        for a, b, c in M.TABLE[cfg]:
            for e in ((a, b), (b, c), (c, a)):
                directed[e] = directed.get(e, 0) + 1
        assert all(v == 1 for v in directed.values()), cfg
        for (a, b) in directed:
            if (b, a) in directed:
                continue
            shared = set(M.EDGES[a]) | set(M.EDGES[b])
            on_face = any(set(M.EDGES[a]) <= set(f) and set(M.EDGES[b]) <= set(f) for f in M.FACES)
            assert on_face, (cfg, a, b, shared)


def test_planar_fields_give_upward_normals():
    """For a linear field f(p) = g.p + b the patch is the plane f = 0 inside
    the cube; every non-degenerate triangle's normal points along g (from
    below to above)."""
    rng = np.random.default_rng(3)
    hit = set()
    for _ in range(20000):
        g = rng.standard_normal(3)
        b = rng.uniform(-1.5, 1.5)
        f = [float(np.dot(g, M.corner_pos(k)) + b) for k in range(8)]
        tris = M.voxel_triangles(f, 0.0, 0, 0, 0, np.float64)
        cfg = sum(1 << k for k in range(8) if f[k] < 0)
        hit.add(cfg)
        for t in tris:
            n = np.cross(t[1].astype(np.float64) - t[0], t[2].astype(np.float64) - t[0])
            if np.linalg.norm(n) > 1e-6:
                assert np.dot(n, g) > 0, (cfg, f)
    assert len(hit) > 100  # many configurations are planar cuts


def test_shared_faces_are_cut_alike():
    """For two voxels sharing a face, the segments each cuts on that face
    are the same set of corner pairs (so the surface has no cracks): checked
    for every sign pattern of the 12 corners of two voxels stacked along x."""
    def face_segments(cfg, face):
        """Directed triangle edges of the patch that lie on `face` (boundary
        segments: the patch's interior edges are not on a face)."""
        segs = set()
        for a, b, c in M.TABLE[cfg]:
            for e1, e2 in ((a, b), (b, c), (c, a)):
                if set(M.EDGES[e1]) <= set(face) and set(M.EDGES[e2]) <= set(face):
                    segs.add((M.EDGES[e1], M.EDGES[e2]))
        return segs

    def shared_ids(edge):
        """A face edge as its two (y, z) corner positions, the same for both voxels."""
        return tuple(sorted(((k >> 1) & 3) for k in edge))
    right_face = [1, 3, 5, 7]   # x = 1 face of the left voxel
    left_face = [0, 2, 4, 6]    # x = 0 face of the right voxel
    for pattern in itertools.product((0, 1), repeat=12):
        # left voxel corners: x = 0 plane (4) and shared plane (4); right: shared plane and x = 2 plane
        shared = pattern[4:8]
        lcfg = sum(pattern[i] << k for i, k in enumerate([0, 2, 4, 6])) + sum(shared[i] << k for i, k in enumerate([1, 3, 5, 7]))
        rcfg = sum(shared[i] << k for i, k in enumerate([0, 2, 4, 6])) + sum(pattern[8 + i] << k for i, k in enumerate([1, 3, 5, 7]))
        ls = face_segments(lcfg, right_face)
        rs = face_segments(rcfg, left_face)
        # the same segments, traversed in opposite directions (consistent winding across the face)
        assert {tuple(shared_ids(e) for e in seg) for seg in ls} == {tuple(shared_ids(e) for e in reversed(seg)) for seg in rs}, pattern

"""The mesh voxelizer oracle (oracle/mesh.py) pinned by closed forms: a box
mesh with faces between cell centres fills exactly its cells; convex meshes
agree with their half-space description; overlapping solids take the lowest id."""
import numpy as np

from oracle import mesh as M


def test_box_mesh_fills_its_cells():
    v, f = M.box_mesh((3.5, 2.5, 1.5), (8.5, 6.5, 4.5))
    m = M.voxelize((7, 9, 11), M.posed_tris(v, f), np.zeros(12, int))
    want = np.zeros((7, 9, 11), np.uint8)
    want[2:5, 3:7, 4:9] = 1
    assert np.array_equal(m, want)


def _halfspace(v, f, p):
    """Inside a closed convex mesh: behind every face plane (outward normals)."""
    c = v.mean(0)
    inside = np.ones(len(p), bool)
    for a, b, d in f:
        n = np.cross(v[b] - v[a], v[d] - v[a])
        if np.dot(n, v[a] - c) < 0:
            n = -n
        inside &= (p - v[a]) @ n < 0
    return inside


def test_convex_mesh_matches_halfspaces():
    v, f = M.icosphere(4.3, 1)
    R = np.array([[0.36, 0.48, -0.8], [-0.8, 0.6, 0.0], [0.48, 0.64, 0.6]])
    t = np.array([7.2, 6.9, 6.4])
    m = M.voxelize((13, 14, 15), M.posed_tris(v, f, R, t), np.zeros(len(f), int))
    z, y, x = np.mgrid[0:13, 0:14, 0:15]
    p = np.stack([x.ravel(), y.ravel() + M.DY, z.ravel() + M.DZ], 1).astype(np.float64)
    want = _halfspace(v @ R.T + t, f, p).reshape(13, 14, 15)
    assert want.sum() > 100
    assert np.array_equal(m != 0, want)


def test_lowest_solid_wins_and_merge():
    va, fa = M.box_mesh((1.5, 1.5, 1.5), (5.5, 5.5, 5.5))
    vb, fb = M.box_mesh((3.5, 3.5, 3.5), (7.5, 7.5, 7.5))
    tris = np.concatenate([M.posed_tris(vb, fb), M.posed_tris(va, fa)])
    sid = np.array([1] * 12 + [0] * 12)
    m = M.voxelize((9, 9, 9), tris, sid)
    assert m[4, 4, 4] == 1 and m[6, 6, 6] == 2 and m[2, 2, 2] == 1 and m[8, 8, 8] == 0
    base = np.zeros((9, 9, 9), np.uint8)
    base[4, 4, 4] = 7
    mm = M.voxelize((9, 9, 9), M.posed_tris(va, fa), [0] * 12, mask=base, merge=True)
    assert mm[4, 4, 4] == 7 and mm[2, 2, 2] == 1 and mm[8, 8, 8] == 0


def test_hollow_shell_is_outer_minus_inner():
    """A non-convex body: an outer box with an inner box as a cavity (one
    solid); parity leaves the cavity empty."""
    vo, fo = M.box_mesh((1.5, 1.5, 1.5), (9.5, 8.5, 7.5))
    vi, fi = M.box_mesh((3.5, 3.5, 3.5), (6.5, 5.5, 5.5))
    tris = np.concatenate([M.posed_tris(vo, fo), M.posed_tris(vi, fi[:, ::-1])])
    m = M.voxelize((10, 11, 12), tris, np.zeros(24, int))
    want = np.zeros((10, 11, 12), np.uint8)
    want[2:8, 2:9, 2:10] = 1
    want[4:6, 4:6, 4:7] = 0
    assert np.array_equal(m, want)

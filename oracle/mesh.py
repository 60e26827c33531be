"""CPU restatement of the triangle-mesh voxelizer (test infrastructure only).

A cell centre (x, y, z) is inside when the ray from (x, y + 1.2345679e-4,
z + 2.3456789e-4) along +x crosses its solid's triangles an odd number of
times; a triangle is crossed when the ray lies strictly inside its yz
projection (edge functions all of one sign), at the x of its barycentric
weights; mask = lowest solid with odd parity + 1. FP32 in the kernel's
operation order. No reference counterpart to follow line by line (the
reference ray-casts a Warp BVH, envs/lbm3d/lbm_func_3d.py:254-316); the
tests pin it with closed forms: box-shaped meshes, and convex meshes against
their half-space description.
"""
from __future__ import annotations

import numpy as np

F = np.float32
DY, DZ = F(1.2345679e-4), F(2.3456789e-4)


def voxelize(shape_zyx, tris, solids, box=None, mask=None, merge=False):
    """Mask [nz, ny, nx] uint8 over the box (x0, x1, y0, y1, z0, z1)."""
    nz, ny, nx = shape_zyx
    m = np.zeros(shape_zyx, np.uint8) if mask is None else mask.copy()
    if box is None:
        box = (0, nx, 0, ny, 0, nz)
    T = np.asarray(tris, F).reshape(-1, 9)
    S = np.asarray(solids, np.int64).reshape(-1)
    x0, x1, y0, y1, z0, z1 = box
    for z in range(z0, z1):
        for y in range(y0, y1):
            py, pz = F(y) + DY, F(z) + DZ
            ay, az, by, bz, cy, cz = (T[:, 1] - py, T[:, 2] - pz, T[:, 4] - py, T[:, 5] - pz,
                                      T[:, 7] - py, T[:, 8] - pz)
            w0 = by * cz - bz * cy
            w1 = cy * az - cz * ay
            w2 = ay * bz - az * by
            hit = ((w0 > 0) & (w1 > 0) & (w2 > 0)) | ((w0 < 0) & (w1 < 0) & (w2 < 0))
            xc = (w0[hit] * T[hit, 0] + w1[hit] * T[hit, 3] + w2[hit] * T[hit, 6]) / (w0[hit] + w1[hit] + w2[hit])
            sc = S[hit]
            for x in range(x0, x1):
                bits = 0
                for xv, s in zip(xc, sc):
                    if xv > F(x):
                        bits ^= 1 << int(s)
                v = (bits & -bits).bit_length() if bits else 0
                if not merge:
                    m[z, y, x] = v
                elif v and m[z, y, x] == 0:
                    m[z, y, x] = v
    return m


def box_mesh(lo, hi):
    """Closed axis-aligned box: 8 vertices, 12 triangles (outward normals)."""
    (x0, y0, z0), (x1, y1, z1) = lo, hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float64)
    f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [2, 3, 7], [2, 7, 6], [1, 2, 6], [1, 6, 5], [0, 4, 7], [0, 7, 3]])
    return v, f


def icosphere(radius, subdivisions=1):
    """Closed convex polyhedron inscribed in a sphere (subdivided icosahedron)."""
    t = (1 + 5 ** 0.5) / 2
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    v = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(subdivisions):
        cache, nf = {}, []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                cache[k] = len(v) - 1
            return cache[k]
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    return np.array(v) * radius, np.array(f)


def posed_tris(verts, faces, R=None, t=(0, 0, 0)):
    v = np.asarray(verts, np.float64)
    if R is not None:
        v = v @ np.asarray(R, np.float64).T
    v = v + np.asarray(t, np.float64)
    return v[np.asarray(faces)].reshape(-1, 9).astype(F)

"""GPU triangle-mesh voxelizer against the mesh oracle, bitwise, and mesh
bodies through LBM_Solver3D.create_solid_from_mesh."""
import numpy as np
import pytest
import torch

from oracle import mesh as M
from openhomelbm_amd.lattice import Lattice
from openhomelbm_amd.lbm3d import LBM_Solver3D, quat_to_matrix

pytestmark = pytest.mark.gpu
SHAPE = (20, 22, 26)  # nz, ny, nx


def _mask(lat):
    torch.cuda.synchronize()
    return lat.mask.view(SHAPE).cpu().numpy()


def test_box_and_rotated_convex_mesh_bit_exact():
    lat = Lattice(27, SHAPE[2], SHAPE[1], SHAPE[0], nsolid=2)
    v, f = M.box_mesh((3.5, 2.5, 1.5), (8.5, 6.5, 4.5))
    tris = M.posed_tris(v, f)
    lat.voxelize_mesh(tris, [0] * len(f), box=(0, SHAPE[2], 0, SHAPE[1], 0, SHAPE[0]))
    assert np.array_equal(_mask(lat), M.voxelize(SHAPE, tris, [0] * len(f)))
    vi, fi = M.icosphere(6.1, 2)
    R = quat_to_matrix([0.8, 0.2, -0.4, 0.3])
    tris = M.posed_tris(vi, fi, R, (13.3, 10.8, 9.7))
    lat.voxelize_mesh(tris, [1] * len(fi))
    want = M.voxelize(SHAPE, tris, [1] * len(fi), box=lat.mask_box)
    assert np.array_equal(_mask(lat), want) and (want == 2).sum() > 500
    assert not lat.mesh_overflowed()


def test_overlapping_solids_and_merge_bit_exact():
    lat = Lattice(27, SHAPE[2], SHAPE[1], SHAPE[0], nsolid=3)
    va, fa = M.box_mesh((2.5, 2.5, 2.5), (9.5, 9.5, 9.5))
    vb, fb = M.icosphere(4.2, 1)
    tris = np.concatenate([M.posed_tris(vb, fb, None, (9.2, 9.1, 8.8)), M.posed_tris(va, fa)])
    sid = [2] * len(fb) + [0] * len(fa)
    full = (0, SHAPE[2], 0, SHAPE[1], 0, SHAPE[0])
    lat.voxelize_mesh(tris, sid, box=full)
    want = M.voxelize(SHAPE, tris, sid)
    assert np.array_equal(_mask(lat), want) and (want == 1).any() and (want == 3).any()
    # capsule mask first, then a mesh merged into it
    lat2 = Lattice(27, SHAPE[2], SHAPE[1], SHAPE[0], nsolid=3)
    lat2.voxelize(np.array([[18, 14, 12, 21, 16, 12, 2.5, 1]], np.float32), box=full)
    base = _mask(lat2).copy()
    lat2.voxelize_mesh(tris, sid, merge=True)
    want = M.voxelize(SHAPE, tris, sid, mask=base, merge=True)
    assert np.array_equal(_mask(lat2), want) and (base > 0).any() and (want != base).any()


def test_create_solid_from_mesh_voxelizes_and_feels_drag():
    nx, ny, nz = 48, 28, 26
    s = LBM_Solver3D(nx, ny, nz, solid_num=1)
    v, f = M.icosphere(1.0, 2)  # unit sphere in body units
    info = s.create_solid_from_mesh(0, (v, f), lbm_position=(16.0, 14.0, 13.0), lbm_scale=5.0 / nx,
                                    init_quaternion=(0.9, 0.1, 0.3, -0.2))
    assert info["scale"] == 5.0
    fl = s.flows[0]
    fl.vis_shear = 0.05
    fl.bc_type = [0, 1, 1, 1, 1, 1]
    fl.bc_value = [[0.04, 0, 0]] + [[0, 0, 0]] * 5
    fl.lattice.init_uniform((0.04, 0.0, 0.0))
    for _ in range(60):
        s.step()
    torch.cuda.synchronize()
    R = quat_to_matrix((0.9, 0.1, 0.3, -0.2))
    tris = M.posed_tris(v * 5.0, f, R, (16.0, 14.0, 13.0))
    got = fl.lattice.mask.view(nz, ny, nx).cpu().numpy()
    assert np.array_equal(got, M.voxelize((nz, ny, nx), tris, [0] * len(f)))
    F = fl.solid_force.numpy()[0]
    assert np.isfinite(F).all() and F[0] > 0 and abs(F[1]) < 0.2 * F[0] and abs(F[2]) < 0.2 * F[0]


def test_capsule_and_mesh_bodies_in_one_world_and_moving():
    """A capsule (solid 0) and a mesh (solid 1) in one world: the mask is the
    capsule voxelization with the mesh merged into its empty cells; after the
    mesh body moves, its old cells are cleared."""
    from oracle import oracle as orc

    nx, ny, nz = 40, 26, 24
    s = LBM_Solver3D(nx, ny, nz, solid_num=2)
    fl = s.flows[0]
    s.create_solid_capsule(0, (-0.1, 0.0, 0.0), (0.1, 0.0, 0.0), 0.08, lbm_position=(12.0, 13.0, 12.0),
                           lbm_scale=1.0)
    v, f = M.icosphere(1.0, 1)
    s.create_solid_from_mesh(1, (v, f), lbm_position=(26.3, 12.6, 11.7), lbm_scale=4.5 / nx)

    def want():
        env = orc.Env(27, nx, ny, nz)
        base = orc.voxelize(env, fl.capsules)
        pos = fl.solid_position.numpy()[1]
        R = quat_to_matrix(fl.solid_quaternion.numpy()[1])
        tris = M.posed_tris(v * 4.5, f, R, pos)
        return M.voxelize((nz, ny, nx), tris, [1] * len(f), mask=base, merge=True)

    for k in range(3):
        if k == 2:
            s.update_solids_batch(0, [1], np.array([[0.1, -0.05, 0.0]]), np.array([[0.95, 0.0, 0.31, 0.0]]))
        s.step()
        torch.cuda.synchronize()
        got = fl.lattice.mask.view(nz, ny, nx).cpu().numpy()
        w = want()
        assert np.array_equal(got, w), k
        assert set(np.unique(w)) == {0, 1, 2}

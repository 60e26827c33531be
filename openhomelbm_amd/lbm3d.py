"""D3Q27 solver API of the reference (envs/lbm3d/lbm_core_3d.py HomeFlow3D and
envs/lbm3d/lbm_solver_3d.py LBM_Solver3D) over the HIP library.

Same class names, constructor signatures, methods and field names; fields
are LatticeField / HostArray handles with .numpy() / .assign(). Differences,
all deliberate:

* Bodies are closed triangle meshes (create_solid_from_mesh) or analytic
  capsules and spheres (create_solid_capsule, create_solid_sphere) in the body
  frame, scaled by lbm_scale * nx and posed by position + quaternion
  (w, x, y, z). When a pose or shape changes they are voxelized into an
  integer mask (meshes by +x ray parity) and the wall rule runs on mask links
  (mask(x) != mask(x-c_i)) instead of BVH ray queries (get_cutcell_multi_3d,
  lbm_func_3d.py:254-316). With HomeFlow3D.cut_cell = True (or
  LBM_Solver3D(..., cut_cell=True)), capsule and sphere bodies take the
  reference's own rule instead: a link is cut where the ray from the cell
  along -c_i meets a capsule surface at t in [0, 1] (closed form in place of
  the BVH), with u_p and the lever arm at the crossing x - c_i t.
* solid_force / solid_torque come from a deterministic FP64 reduction.
* step() is asynchronous; reading a field synchronizes.
"""
from __future__ import annotations

import math
from typing import Any, Dict, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native as N
from .arrays import ForceField, HostArray, LatticeField, as_numpy
from .lattice import Lattice, capsule_box, step_worlds, union_box


def _tri_box(tris: np.ndarray, L: Lattice):
    """Half-open cell box around posed triangles (one cell of padding)."""
    v = np.asarray(tris, np.float64).reshape(-1, 3)
    lo, hi = v.min(0), v.max(0)
    box = []
    for a, n in enumerate((L.nx, L.ny, L.nz)):
        b0 = max(0, int(math.floor(lo[a])) - 1)
        box += [b0, max(b0, min(n, int(math.ceil(hi[a])) + 2))]
    return tuple(box)

ML_FLUID, ML_SOLID, ML_WALL = 0, 1, 2

_E = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0),
      (1, 0, 1), (-1, 0, -1), (0, 1, 1), (0, -1, -1), (1, -1, 0), (-1, 1, 0), (1, 0, -1), (-1, 0, 1), (0, 1, -1),
      (0, -1, 1), (1, 1, 1), (-1, -1, -1), (1, 1, -1), (-1, -1, 1), (1, -1, 1), (-1, 1, -1), (-1, 1, 1),
      (1, -1, -1)]


def quat_to_matrix(q) -> np.ndarray:
    """Rotation matrix of a (w, x, y, z) quaternion (normalised first)."""
    w, x, y, z = (float(v) for v in q)
    n = math.sqrt(w * w + x * x + y * y + z * z)
    if n == 0.0:
        raise ValueError("zero quaternion")
    w, x, y, z = w / n, x / n, y / n, z / n
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])



def first_step_last_pose(now_row: np.ndarray, linear_v, angle_v) -> np.ndarray:
    """Previous-step pose row (p, R) of a body whose transforms are not
    initialized yet, chosen so that the wall velocity the kernels take from
    the pose pair (the body point's position now minus the same body-frame
    point's position a step earlier) is the reference's fallback
    u_p = linear_v + angle_v x (h - p) (lbm_func_3d.py:459-474):
    p_last = p - v and R_last = R - [w]x R, so R q + p - (R_last q + p_last)
    = w x (R q) + v with R q = h - p. With zero velocities the body is at rest
    (R_last = R, p_last = p)."""
    row = np.array(now_row, np.float64, copy=True)
    v = np.asarray(linear_v, np.float64).reshape(3)
    w = np.asarray(angle_v, np.float64).reshape(3)
    R = row[3:12].reshape(3, 3)
    W = np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])
    row[0:3] = row[0:3] - v
    row[3:12] = (R - W @ R).reshape(9)
    return row

class HomeFlow3D:
    """D3Q27 state of one world (envs/lbm3d/lbm_core_3d.py:45-227)."""

    def __init__(self, dtype: str = "f32", device: Optional[torch.device] = None):
        self._dtype = dtype
        self._device = device
        self._lat: Optional[Lattice] = None
        self._images: dict = {}  # render buffers written by openhomelbm_amd.render

    def Initialize(self, nx: int, ny: int, nz: int, n_objects: int = 1) -> None:
        """Allocate the world: rho = 1, u = 0, S = 0 (InitFlow3D,
        lbm_func_3d.py:22-83), bc_type all 1, vis_shear 0.1."""
        if n_objects < 0 or n_objects > N.HL_MAX_SOLIDS:
            raise ValueError(f"n_objects must be in [0, {N.HL_MAX_SOLIDS}]")
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.time_step = 1.0
        self.grid_length = 1.0 / nx
        self.velocity_scale = self.time_step / self.grid_length
        self._lat = Lattice(27, self.nx, self.ny, self.nz, dtype=self._dtype, nsolid=n_objects, device=self._device)
        self._uploaded = None
        L = self._lat
        L.nu = 0.1
        L.bc_type = [1] * 6
        L.init_uniform((0.0, 0.0, 0.0))
        self.n_objects = n_objects
        self.n_dirs = 27
        self.lat_e = np.array(_E, np.float32)
        self.lat_inv = np.array([_E.index((-a, -b, -c)) for a, b, c in _E], np.int32)
        w = [8 / 27 if sum(map(abs, e)) == 0 else 2 / 27 if sum(map(abs, e)) == 1 else
             1 / 54 if sum(map(abs, e)) == 2 else 1 / 216 for e in _E]
        self.lat_w = np.array(w, np.float32)
        n = max(n_objects, 1)
        dirty = self._mark_dirty
        self._solid = {
            "solid_position": HostArray(np.zeros((n, 3)), on_change=dirty),
            "solid_quaternion": HostArray(np.tile([1.0, 0.0, 0.0, 0.0], (n, 1)), on_change=dirty),
            "solid_bound_radius": HostArray(np.zeros(n)),
            "mass": HostArray(np.zeros(n)),
            "MoI": HostArray(np.zeros(n)),
            "mesh_scale_sizes": HostArray(np.ones((n, 3))),
            "mesh_transforms_initialized": HostArray(np.zeros(n, np.int32), dtype=np.int32),
            # lattice units per step: the wall velocity of a body whose
            # transforms are not initialized yet (lbm_core_3d.py:105-106, 202-203)
            "linear_v": HostArray(np.zeros((n, 3))),
            "angle_v": HostArray(np.zeros((n, 3))),
        }
        self._shapes: Dict[int, Tuple[np.ndarray, np.ndarray, float]] = {}  # body-frame capsules (lattice units)
        self._meshes: Dict[int, Tuple[np.ndarray, np.ndarray]] = {}  # body-frame meshes (lattice units)
        self._last: Optional[np.ndarray] = None  # pose table of the previous step
        self._uploaded: Optional[np.ndarray] = None  # the [now | last] table the lattice holds
        self.capsules = np.zeros((0, 8), np.float32)  # world-frame capsules of the last voxelization
        self._caps_dev: Optional[torch.Tensor] = None  # the same table on the device (capsule cut links)
        self._caps_box = (0, 0, 0, 0, 0, 0)
        # cut_cell: capsule and sphere bodies cut links where the link crosses
        # their surface, with the wall rule at the crossing (the reference's
        # ray query, lbm_func_3d.py:254-316, in closed form); False: the mask
        # rule at link midpoints. Mesh bodies always take the mask rule.
        self.cut_cell = False
        self._dirty = True
        self._forces = torch.zeros((1, n, 6), dtype=torch.float64, device=L.device)

    def _mark_dirty(self):
        self._dirty = True

    @property
    def lattice(self) -> Lattice:
        if self._lat is None:
            raise RuntimeError("HomeFlow3D.Initialize(nx, ny, nz, n_objects) has not been called")
        return self._lat

    rho = property(lambda s: LatticeField(s.lattice, (0,)))
    u = property(lambda s: LatticeField(s.lattice, (1, 2, 3)))
    Sxx = property(lambda s: LatticeField(s.lattice, (4,)))
    Syy = property(lambda s: LatticeField(s.lattice, (5,)))
    Szz = property(lambda s: LatticeField(s.lattice, (6,)))
    Sxy = property(lambda s: LatticeField(s.lattice, (7,)))
    Sxz = property(lambda s: LatticeField(s.lattice, (8,)))
    Syz = property(lambda s: LatticeField(s.lattice, (9,)))
    rho_post = property(lambda s: LatticeField(s.lattice, (0,), "post"))
    u_post = property(lambda s: LatticeField(s.lattice, (1, 2, 3), "post"))
    Sxx_post = property(lambda s: LatticeField(s.lattice, (4,), "post"))
    Syy_post = property(lambda s: LatticeField(s.lattice, (5,), "post"))
    Szz_post = property(lambda s: LatticeField(s.lattice, (6,), "post"))
    Sxy_post = property(lambda s: LatticeField(s.lattice, (7,), "post"))
    Sxz_post = property(lambda s: LatticeField(s.lattice, (8,), "post"))
    Syz_post = property(lambda s: LatticeField(s.lattice, (9,), "post"))
    forcex = property(lambda s: ForceField(s.lattice, 0))
    forcey = property(lambda s: ForceField(s.lattice, 1))
    forcez = property(lambda s: ForceField(s.lattice, 2))

    def _image(self, name: str, shape) -> HostArray:
        img = self._images.get(name)
        if img is None:
            return HostArray(np.zeros(shape, np.float32))
        return HostArray(img.cpu().numpy())

    # render buffers written by openhomelbm_amd.render (zeros before), in the
    # reference's index order (lbm_core_3d.py:149-222)
    u_img_xy = property(lambda s: s._image("u_img_xy", (s.nx, s.ny)))
    u_img_xz = property(lambda s: s._image("u_img_xz", (s.ny, s.nz)))
    u_img_xz_front = property(lambda s: s._image("u_img_xz_front", (s.nx, s.nz)))

    @property
    def flag(self) -> HostArray:
        """ML_WALL on every face cell (InitBoundary3D, lbm_func_3d.py:13-19)."""
        f = np.zeros((self.nx, self.ny, self.nz), np.int32)
        f[0], f[-1], f[:, 0], f[:, -1], f[:, :, 0], f[:, :, -1] = (ML_WALL,) * 6
        return HostArray(f, dtype=np.int32)

    @property
    def vis_shear(self) -> float:
        return float(self.lattice.nu)

    @vis_shear.setter
    def vis_shear(self, v) -> None:
        v = float(v)
        if not v > 0.0 or not math.isfinite(v):
            raise ValueError(f"vis_shear must be > 0, got {v}")
        self.lattice.nu = v

    @property
    def bc_type(self):
        return tuple(int(t) for t in self.lattice.bc_type[:6])

    @bc_type.setter
    def bc_type(self, v) -> None:
        t = [int(x) for x in as_numpy(v).reshape(-1)]
        if len(t) != 6:
            raise ValueError(f"bc_type needs 6 entries (left, right, top, bottom, front, back), got {len(t)}")
        if any(x not in (0, 1, 2) for x in t):
            raise ValueError("bc_type entries must be 0 (velocity), 1 (zero gradient) or 2 (free slip)")
        self.lattice.bc_type = t

    @property
    def bc_value(self) -> HostArray:
        return HostArray(np.array(self.lattice.bc_value[:6]))

    @bc_value.setter
    def bc_value(self, v) -> None:
        a = as_numpy(v, np.float64)
        if a.shape != (6, 3):
            raise ValueError(f"bc_value needs shape (6, 3), got {a.shape}")
        self.lattice.bc_value = [[float(x) for x in a[k]] for k in range(6)]

    def __getattr__(self, name):
        solid = self.__dict__.get("_solid")
        if solid is not None and name in solid:
            return solid[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        solid = self.__dict__.get("_solid")
        if solid is not None and name in solid:
            old = solid[name]
            solid[name] = HostArray(value, dtype=old.dtype, shape=old.shape, on_change=old._on_change)
            self._dirty = True
            return
        object.__setattr__(self, name, value)

    @property
    def solid_force(self) -> HostArray:
        return HostArray(self._forces[0, : self.n_objects, 0:3].cpu().numpy())

    @property
    def solid_torque(self) -> HostArray:
        return HostArray(self._forces[0, : self.n_objects, 3:6].cpu().numpy())

    # ------------------------------------------------------------ bodies
    def set_capsule(self, solid_id: int, a, b, radius: float) -> None:
        """Body-frame capsule (lattice units) of solid solid_id."""
        if solid_id < 0 or solid_id >= self.n_objects:
            raise ValueError(f"solid_id {solid_id} out of range [0, {self.n_objects})")
        if not radius > 0:
            raise ValueError("radius must be > 0")
        a = np.asarray(a, np.float64).reshape(3)
        b = np.asarray(b, np.float64).reshape(3)
        self._shapes[solid_id] = (a, b, float(radius))
        self._solid["solid_bound_radius"][solid_id] = max(np.linalg.norm(a), np.linalg.norm(b)) + radius
        self._dirty = True

    def set_mesh(self, solid_id: int, vertices, faces) -> None:
        """Body-frame closed triangle mesh (lattice units) of solid solid_id."""
        if solid_id < 0 or solid_id >= self.n_objects:
            raise ValueError(f"solid_id {solid_id} out of range [0, {self.n_objects})")
        v = np.asarray(vertices, np.float64).reshape(-1, 3)
        f = np.asarray(faces, np.int64).reshape(-1, 3)
        if f.size and (f.min() < 0 or f.max() >= len(v)):
            raise ValueError("face indices out of range")
        self._meshes[solid_id] = (v, f)
        self._shapes.pop(solid_id, None)
        self._solid["solid_bound_radius"][solid_id] = float(np.linalg.norm(v, axis=1).max()) if len(v) else 0.0
        self._dirty = True

    def _pose_table(self) -> np.ndarray:
        n = max(self.n_objects, 1)
        P = np.zeros((n, 12))
        pos, quat = self._solid["solid_position"].numpy(), self._solid["solid_quaternion"].numpy()
        for k in range(n):
            P[k, 0:3] = pos[k]
            P[k, 3:12] = quat_to_matrix(quat[k]).reshape(9)
        return P

    def _prepare(self, stream=None) -> None:
        """Pose table (current | previous step) and, when poses or shapes
        changed, the mask from the posed capsules."""
        if self.n_objects == 0:
            return
        L = self.lattice
        now = self._pose_table()
        init = self._solid["mesh_transforms_initialized"].numpy()
        last = now.copy() if self._last is None else self._last.copy()
        lv, av = self._solid["linear_v"].numpy(), self._solid["angle_v"].numpy()
        for k in range(self.n_objects):
            if not init[k]:
                last[k] = first_step_last_pose(now[k], lv[k], av[k])
        init[: self.n_objects] = 1
        self._solid["mesh_transforms_initialized"]._a[...] = init
        table = np.concatenate([now, last], axis=1)[: self.n_objects].astype(np.float32)
        # the lattice's table is this object's alone: an unchanged table (bodies
        # at rest) is not copied again
        if self._uploaded is None or not np.array_equal(table, self._uploaded):
            L.set_poses(table, stream=stream)
            self._uploaded = table
        self._last = now
        if not self._dirty:
            return
        caps = []
        for k, (a, b, r) in sorted(self._shapes.items()):
            R, p = now[k, 3:12].reshape(3, 3), now[k, 0:3]
            caps.append(list(R @ a + p) + list(R @ b + p) + [r, k])
        self.capsules = np.array(caps, np.float32).reshape(-1, 8)
        self._caps_dev = torch.from_numpy(self.capsules).to(L.device) if caps else None
        self._caps_box = capsule_box(self.capsules, (self.nx, self.ny, self.nz)) if caps else (0, 0, 0, 0, 0, 0)
        tris, tsid = [], []
        for k, (v, f) in sorted(self._meshes.items()):
            R, p = now[k, 3:12].reshape(3, 3), now[k, 0:3]
            w = v @ R.T + p
            tris.append(w[f].reshape(-1, 9))
            tsid.append(np.full(len(f), k, np.int32))
        if caps and tris:
            T = np.concatenate(tris).astype(np.float32)
            box = union_box(capsule_box(self.capsules, (self.nx, self.ny, self.nz)), _tri_box(T, L))
            L.voxelize(self.capsules, box=box, stream=stream)
            L.voxelize_mesh(T, np.concatenate(tsid), merge=True, stream=stream)
        elif caps:
            L.voxelize(self.capsules, stream=stream)
        elif tris:
            L.voxelize_mesh(np.concatenate(tris), np.concatenate(tsid), stream=stream)
        self._dirty = False

    def _uses_capsule_cuts(self) -> bool:
        if not self.cut_cell or self.n_objects == 0:
            return False
        if self._meshes:
            raise ValueError("cut_cell needs every body to be a capsule or a sphere (mesh bodies take the mask "
                             "rule: set cut_cell = False)")
        return self._caps_dev is not None

    def step(self, stream=None) -> None:
        self._prepare(stream)
        self._step_prepared(stream)

    def _step_prepared(self, stream=None) -> None:
        """The lattice step after _prepare (pose table and mask current)."""
        fh = self._forces if self.n_objects > 0 else None
        if self._uses_capsule_cuts():
            self.lattice.step_capsules(1, self._caps_dev, self._caps_box, force_history=fh, stream=stream)
        else:
            self.lattice.step(1, N.HL_KERNEL_AUTO, force_history=fh, stream=stream)

    def reset(self) -> None:
        """ResetSingleWorldFlow3D (lbm_func_3d.py:87-140): rho = 1, u = 0, S = 0
        in both states; the transform history restarts."""
        L = self.lattice
        L.init_uniform((0.0, 0.0, 0.0))
        self._solid["mesh_transforms_initialized"]._a[...] = 0
        self._last = None


class LBM_Solver3D:
    """Worlds of one size stepped together (envs/lbm3d/lbm_solver_3d.py:19-420)."""

    def __init__(self, nx: int, ny: int, nz: int, solid_num: int = 1, nworld: int = 1,
                 device: Optional[Any] = None, dtype: str = "f32", cut_cell: bool = False) -> None:
        if nworld < 1:
            raise ValueError("nworld must be >= 1")
        self.nx, self.ny, self.nz = nx, ny, nz
        self.solid_num = solid_num
        self.nworld = nworld
        self.captured = False
        self.captured_graph = None
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.flows = [HomeFlow3D(dtype=dtype, device=self.device) for _ in range(nworld)]
        for flow in self.flows:
            flow.Initialize(nx, ny, nz, n_objects=solid_num)
            flow.cut_cell = cut_cell
        self.flows_wp = self.flows
        self.meshes = [None] * solid_num
        self.mesh_wps = [None] * solid_num
        self.mujoco_mappings: Dict[int, dict] = {}
        self.mujoco_origins_wp = None
        self.lbm_origins_wp = None
        self.scales_wp = None
        self.solid_ids_wp = None
        self.solid_id_to_index: Dict[int, int] = {}

    def step(self) -> None:
        """Advance every world one step (asynchronous). The reference launches
        every world in one grid (lbm_solver_3d.py:104-108); here the worlds go
        to the library in one call (hl_step_worlds: up to six worlds per tiled
        launch, each world bitwise as if stepped alone). Worlds with capsule
        cut links step one by one."""
        flows = self.flows
        for flow in flows:
            flow._prepare()
        if len(flows) > 1 and not any(flow._uses_capsule_cuts() for flow in flows):
            step_worlds([flow.lattice for flow in flows], 1, N.HL_KERNEL_AUTO,
                        [flow._forces if flow.n_objects > 0 else None for flow in flows])
        else:
            for flow in flows:
                flow._step_prepared()
        self.captured = True

    def _check_solid(self, solid_id):
        if solid_id < 0 or solid_id >= self.solid_num:
            raise ValueError(f"solid_id {solid_id} out of range [0, {self.solid_num})")

    def create_solid_from_mesh(self, solid_id, mesh, lbm_position=None, lbm_scale=0.15, init_quaternion=(1, 0, 0, 0),
                               mujoco_origin=None) -> Dict[str, Any]:
        """The same closed triangle mesh in every world (lbm_solver_3d.py:129-242):
        mesh has .vertices [n, 3] and .faces [m, 3] (a trimesh.Trimesh or any
        object with those arrays, or a (vertices, faces) pair) in its body frame,
        not recentred; scaled by lbm_scale * nx, its origin placed at
        lbm_position (the domain centre by default) with init_quaternion
        (w, x, y, z). The posed mesh is voxelized into each world's mask (the
        reference ray-casts a BVH per link instead)."""
        self._check_solid(solid_id)
        if isinstance(mesh, (tuple, list)) and len(mesh) == 2:
            verts, faces = mesh
        else:
            verts, faces = mesh.vertices, mesh.faces
        s = lbm_scale * self.nx
        scaled = np.asarray(verts, np.float64).reshape(-1, 3) * s
        if lbm_position is None:
            lbm_position = (self.nx * 0.5, self.ny * 0.5, self.nz * 0.5)
        for flow in self.flows:
            flow.set_mesh(solid_id, scaled, faces)
            flow.solid_position[solid_id] = lbm_position
            flow.solid_quaternion[solid_id] = init_quaternion
        if mujoco_origin is None:
            mujoco_origin = np.zeros(3, np.float32)
        self.mujoco_mappings[solid_id] = {"mujoco_origin": np.array(mujoco_origin, np.float32),
                                          "lbm_origin": np.array(lbm_position, np.float32), "scale": s}
        self.meshes[solid_id] = mesh
        return {"mesh": mesh, "vertices": scaled, "scale": s, "lbm_position": lbm_position,
                "quaternion": init_quaternion}

    def create_solid_capsule(self, solid_id: int, p0: Sequence[float], p1: Sequence[float], radius: float,
                             lbm_position: Optional[Sequence[float]] = None, lbm_scale: float = 0.15,
                             init_quaternion: Sequence[float] = (1, 0, 0, 0),
                             mujoco_origin: Optional[np.ndarray] = None) -> Dict[str, Any]:
        """Capsule body in every world: segment p0-p1 and radius in body units
        (scaled by lbm_scale * nx), its body origin placed at lbm_position (the
        domain centre by default) with init_quaternion (w, x, y, z). Mirrors
        create_solid_from_mesh (lbm_solver_3d.py:129-242), mapping included."""
        self._check_solid(solid_id)
        s = lbm_scale * self.nx
        if lbm_position is None:
            lbm_position = (self.nx * 0.5, self.ny * 0.5, self.nz * 0.5)
        a = np.asarray(p0, np.float64) * s
        b = np.asarray(p1, np.float64) * s
        for flow in self.flows:
            flow.set_capsule(solid_id, a, b, radius * s)
            flow.solid_position[solid_id] = lbm_position
            flow.solid_quaternion[solid_id] = init_quaternion
        if mujoco_origin is None:
            mujoco_origin = np.zeros(3, np.float32)
        self.mujoco_mappings[solid_id] = {"mujoco_origin": np.array(mujoco_origin, np.float32),
                                          "lbm_origin": np.array(lbm_position, np.float32), "scale": s}
        return {"p0": a, "p1": b, "radius": radius * s, "scale": s, "lbm_position": tuple(lbm_position),
                "quaternion": tuple(init_quaternion)}

    def create_solid_sphere(self, solid_id: int, radius: float, lbm_position: Optional[Sequence[float]] = None,
                            lbm_scale: float = 0.15, mujoco_origin: Optional[np.ndarray] = None) -> Dict[str, Any]:
        """Sphere body (a capsule of zero length)."""
        return self.create_solid_capsule(solid_id, (0, 0, 0), (0, 0, 0), radius, lbm_position, lbm_scale,
                                         (1, 0, 0, 0), mujoco_origin)

    def create_solids_from_mjcf(self, mjcf_path: str, root_position: Optional[Sequence[float]] = None,
                                lbm_scale: float = 0.15, first_solid: int = 0,
                                mesh_search_paths: Optional[Sequence[str]] = None) -> Dict[str, int]:
        """One mesh solid per MJCF body with geometry, in file order from
        solid first_solid (the body meshes the reference's fluid environment
        takes from parse_mjcf, envs/lbm3d/lbm_fluid_env_3d.py:130-131): each
        body's combined geom mesh in its body frame, placed at root_position +
        lbm_scale * nx * (its world position) with its world orientation, and
        mapped from MuJoCo to lattice coordinates as create_solid_from_mesh
        maps it. Returns {body name: solid id}."""
        from . import mjcf

        data = mjcf.parse_mjcf(mjcf_path, list(mesh_search_paths) if mesh_search_paths else None)
        poses = mjcf.get_body_world_poses(data)
        if first_solid + len(data["bodies"]) > self.solid_num:
            raise ValueError(f"{len(data['bodies'])} bodies from solid {first_solid} need solid_num >= "
                             f"{first_solid + len(data['bodies'])}, have {self.solid_num}")
        if root_position is None:
            root_position = (self.nx * 0.5, self.ny * 0.5, self.nz * 0.5)
        s = lbm_scale * self.nx
        ids = {}
        for k, b in enumerate(data["bodies"]):
            p, R = poses[b["name"]]
            sid = first_solid + k
            self.create_solid_from_mesh(sid, b["combined_mesh"], lbm_position=tuple(np.asarray(root_position) + s * p),
                                        lbm_scale=lbm_scale, init_quaternion=tuple(mjcf.matrix_to_quat(R)),
                                        mujoco_origin=p)
            ids[b["name"]] = sid
        return ids

    def update_solids_batch(self, world_idx: int, solid_ids: Sequence[int], positions: np.ndarray,
                            quaternions: np.ndarray) -> None:
        """Poses of several solids of one world (lbm_solver_3d.py:270-306):
        positions map from MuJoCo to lattice coordinates where a mapping exists."""
        flow = self.flows[world_idx]
        positions = np.asarray(positions, np.float64).reshape(-1, 3)
        quaternions = np.asarray(quaternions, np.float64).reshape(-1, 4)
        for i, sid in enumerate(solid_ids):
            self._check_solid(sid)
            if sid in self.mujoco_mappings:
                m = self.mujoco_mappings[sid]
                pos = (positions[i] - m["mujoco_origin"]) * m["scale"] + m["lbm_origin"]
            else:
                pos = positions[i]
            flow.solid_position[sid] = pos
            flow.solid_quaternion[sid] = quaternions[i]

    def get_forces_and_torques(self, world_idx: int,
                               solid_ids: Optional[Sequence[int]] = None) -> Tuple[np.ndarray, np.ndarray]:
        """(N, 3) forces and torques (lbm_solver_3d.py:308-346): divided by the
        mapping scale (torque by its square) where a mapping exists."""
        if solid_ids is None:
            solid_ids = list(range(self.solid_num))
        flow = self.flows[world_idx]
        F = flow._forces[0].cpu().numpy()
        forces = np.zeros((len(solid_ids), 3), np.float32)
        torques = np.zeros((len(solid_ids), 3), np.float32)
        for i, sid in enumerate(solid_ids):
            if sid in self.mujoco_mappings:
                s = self.mujoco_mappings[sid]["scale"]
                forces[i] = F[sid, 0:3] / s
                torques[i] = F[sid, 3:6] / (s * s)
            else:
                forces[i] = F[sid, 0:3]
                torques[i] = F[sid, 3:6]
        return forces, torques

    def reset_world(self, world_idx: int) -> None:
        self.flows[world_idx].reset()

    def reset_worlds(self, world_indices: Sequence[int]) -> None:
        for i in world_indices:
            self.flows[i].reset()

    def finalize_mappings(self, solid_ids: Sequence[int]) -> None:
        """Coordinate-mapping tables (lbm_solver_3d.py:387-420)."""
        origins, lbm, scales = [], [], []
        for idx, sid in enumerate(solid_ids):
            if sid not in self.mujoco_mappings:
                raise ValueError(f"Solid {sid} not initialized")
            m = self.mujoco_mappings[sid]
            origins.append(m["mujoco_origin"])
            lbm.append(m["lbm_origin"])
            scales.append(m["scale"])
            self.solid_id_to_index[sid] = idx
        self.mujoco_origins_wp = np.array(origins, np.float32)
        self.lbm_origins_wp = np.array(lbm, np.float32)
        self.scales_wp = np.array(scales, np.float32)
        self.solid_ids_wp = np.array(solid_ids, np.int32)

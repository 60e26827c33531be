"""D2Q9 solver API of the reference (envs/lbm/lbm_core.py HomeFlow and
envs/lbm/lbm_solver.py LBM_Solver) over the HIP library.

Same class names, constructor signatures, methods and field names. Fields
are LatticeField / HostArray handles with .numpy() / .assign() instead of
Warp arrays; whole-array assignment (flow.bc_value = ..., flow.vis_shear =
...) works as with the reference. Differences, all deliberate:

* Bodies are closed polygons (configure_solid, lbm_core.py:172-282). By
  default they take the reference's own boundary: the segment-link
  predicate and material-point wall rule (get_cutcell, lbm_func.py:545-607;
  hl_step_polygons; several such worlds of an LBM_Solver step in one
  hl_step_polygons_worlds call). LBM_Solver(..., cut_cell=False) (or HomeFlow.cut_cell
  = False) voxelizes them into an integer mask with the reference's even-odd
  inside test and runs the wall rule on mask links (mask(x) != mask(x-c_i)):
  several FP32 worlds then step in one batched launch sequence (faster
  when the bounding circles cover the lattice), but a different boundary
  discretization (on the 512x128 cylinder: Strouhal 3%, mean drag 16%
  apart). Capsule bodies (add_capsule, an extension)
  always use the mask rule.
* Forces and torques are reduced in FP64 in a fixed order and are
  deterministic; solid_forcex / solid_forcey / torque hold them as FP32 like
  the reference's atomics.
* create_solid_from_mujoco takes an MJCF file (or openhomelbm_amd.mjcf.Model)
  instead of a MuJoCo model and data: MuJoCo does not run here, so the
  outline comes from the file's reference configuration (every joint at 0).
* step() is asynchronous like the reference's graph launch; reading a field
  synchronizes.
"""
from __future__ import annotations

import os

import math
from typing import Any, Optional, Sequence

import numpy as np
import torch

from . import _native as N
from .arrays import ForceField, HostArray, LatticeField, as_numpy
from .lattice import Lattice, step_polygons_worlds

ML_FLUID, ML_SOLID, ML_WALL = 0, 1, 2

_W = (4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36)
_CX = (0, 1, -1, 0, 0, 1, -1, -1, 1)
_CY = (0, 0, 0, 1, -1, 1, 1, -1, -1)
_INV = (0, 2, 1, 4, 3, 7, 8, 5, 6)


class HomeFlow:
    """D2Q9 state of one world (envs/lbm/lbm_core.py:18-282).

    Construct, then Initialize(nx, ny, n_objects) to allocate on the GPU."""

    def __init__(self, dtype: str = "f32", device: Optional[torch.device] = None):
        self._dtype = dtype
        self._device = device
        self._lat: Optional[Lattice] = None
        self._images: dict = {}  # render buffers written by openhomelbm_amd.render

    # ------------------------------------------------------------ setup
    def Initialize(self, nx: int, ny: int, n_objects: int = 1) -> None:
        """Allocate the world (lbm_core.py:90-170): rho = 1, u = 0, S = 0,
        bc_type (1, 1, 1, 1), vis_shear 0.01."""
        if n_objects < 0 or n_objects > N.HL_MAX_SOLIDS:
            raise ValueError(f"n_objects must be in [0, {N.HL_MAX_SOLIDS}]")
        self.time_step = 1.0
        self.nx, self.ny = int(nx), int(ny)
        self.grid_length = 1.0 / nx
        self.velocity_scale = self.time_step / self.grid_length
        self._lat = Lattice(9, self.nx, self.ny, dtype=self._dtype, nsolid=n_objects, device=self._device)
        L = self._lat
        L.nu = 0.01
        L.bc_type = [1, 1, 1, 1, 1, 1]
        L.init_uniform((0.0, 0.0))
        self.w_d2q9 = np.array(_W, np.float32)
        self.cx_d2q9 = np.array(_CX, np.float32)
        self.cy_d2q9 = np.array(_CY, np.float32)
        self.indexd2q9Inv_gpu = np.array(_INV, np.int32)
        self.small_size = 5.0
        self.n_objects = n_objects
        n = max(n_objects, 1)
        self.max_segments_per_object = 10
        self._solid_dirty = True
        dirty = self._mark_dirty
        self._solid = {
            "solid_position": HostArray(np.zeros((n, 2)), on_change=dirty),
            "solid_angle": HostArray(np.zeros(n), on_change=dirty),
            "solid_scale": HostArray(np.zeros(n), on_change=dirty),
            "solid_mass_center": HostArray(np.zeros((n, 2)), on_change=dirty),
            "solid_max_radius": HostArray(np.zeros(n)),
            "mass": HostArray(np.zeros(n)),
            "MoI": HostArray(np.zeros(n)),
            "solid_line": HostArray(np.zeros((n, self.max_segments_per_object, 2)), on_change=dirty),
            "solid_line_num": HostArray(np.zeros(n, np.int32), dtype=np.int32, on_change=dirty),
        }
        self._capsules: dict = {}
        self._last_pose: Optional[np.ndarray] = None
        self._uploaded: Optional[np.ndarray] = None  # pose table last copied to the device
        self._at_rest = False  # uploaded table has last == current and the solids are unchanged since
        self._forces = torch.zeros((1, max(n_objects, 1), 6), dtype=torch.float64, device=L.device)
        self._box = (0, 0, 0, 0, 0, 0)
        # cut_cell: polygon bodies take the reference's segment-link cut cells
        # (get_cutcell); False: the mask link rule
        self.cut_cell = True
        self._seg_prev: Optional[np.ndarray] = None  # last step's transformed vertices (zeros before the first)
        self._seg_host = None  # (verts, last, counts, radius2) last uploaded
        self._seg_dev = None

    def _mark_dirty(self):
        self._solid_dirty = True

    @property
    def cut_cell(self) -> bool:
        """Polygon bodies take the reference's segment-link cut cells (True)
        or the mask link rule (False); switching it between steps rebuilds
        the tables the next step needs."""
        return self._cut_cell

    @cut_cell.setter
    def cut_cell(self, value) -> None:
        value = bool(value)
        if getattr(self, "_cut_cell", None) != value:
            self._cut_cell = value
            self._solid_dirty = True
            self._at_rest = False

    @property
    def lattice(self) -> Lattice:
        if self._lat is None:
            raise RuntimeError("HomeFlow.Initialize(nx, ny, n_objects) has not been called")
        return self._lat

    # ------------------------------------------------------------ fields
    rho = property(lambda s: LatticeField(s.lattice, (0,)))
    u = property(lambda s: LatticeField(s.lattice, (1, 2)))
    Sxx = property(lambda s: LatticeField(s.lattice, (3,)))
    Syy = property(lambda s: LatticeField(s.lattice, (4,)))
    Sxy = property(lambda s: LatticeField(s.lattice, (5,)))
    rho_post = property(lambda s: LatticeField(s.lattice, (0,), "post"))
    u_post = property(lambda s: LatticeField(s.lattice, (1, 2), "post"))
    Sxx_post = property(lambda s: LatticeField(s.lattice, (3,), "post"))
    Syy_post = property(lambda s: LatticeField(s.lattice, (4,), "post"))
    Sxy_post = property(lambda s: LatticeField(s.lattice, (5,), "post"))
    forcex = property(lambda s: ForceField(s.lattice, 0))
    forcey = property(lambda s: ForceField(s.lattice, 1))

    @property
    def flag(self) -> HostArray:
        """ML_WALL on every face cell, ML_FLUID inside (InitBoundary, lbm_func.py:17-23)."""
        f = np.zeros((self.nx, self.ny), np.int32)
        f[0, :] = f[-1, :] = f[:, 0] = f[:, -1] = ML_WALL
        return HostArray(f, dtype=np.int32)

    @property
    def u_img(self) -> HostArray:
        """Render buffer [nx, ny]: the last image a render kernel wrote
        (openhomelbm_amd.render), else the velocity magnitude of the present state."""
        img = self._images.get("u_img")
        if img is None:
            from .render import get_u_img
            get_u_img(self)
            img = self._images["u_img"]
        return HostArray(img.cpu().numpy())

    @property
    def u_img_rgb(self) -> HostArray:
        """RGB render buffer [nx, ny, 3] written by get_velocity_rgb (zeros before)."""
        img = self._images.get("u_img_rgb")
        if img is None:
            return HostArray(np.zeros((self.nx, self.ny, 3), np.float32))
        return HostArray(img.cpu().numpy())

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
        return tuple(int(t) for t in self.lattice.bc_type[:4])

    @bc_type.setter
    def bc_type(self, v) -> None:
        t = [int(x) for x in as_numpy(v).reshape(-1)]
        if len(t) != 4:
            raise ValueError(f"bc_type needs 4 entries (left, top, right, bottom), got {len(t)}")
        if any(x not in (0, 1, 2) for x in t):
            raise ValueError("bc_type entries must be 0 (velocity), 1 (zero gradient) or 2 (free slip)")
        self.lattice.bc_type = t + [1, 1]

    @property
    def bc_value(self) -> HostArray:
        return HostArray(np.array([v[:2] for v in self.lattice.bc_value[:4]]), on_change=None)

    @bc_value.setter
    def bc_value(self, v) -> None:
        a = as_numpy(v, np.float64)
        if a.shape != (4, 2):
            raise ValueError(f"bc_value needs shape (4, 2), got {a.shape}")
        self.lattice.bc_value = [[float(a[k, 0]), float(a[k, 1]), 0.0] for k in range(4)] + [[0.0] * 3] * 2

    # solid tables: whole-array assignment replaces the table
    def __getattr__(self, name):
        solid = self.__dict__.get("_solid")
        if solid is not None and name in solid:
            return solid[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        solid = self.__dict__.get("_solid")
        if solid is not None and name in solid:
            old = solid[name]
            solid[name] = HostArray(value, dtype=old.dtype, on_change=old._on_change)
            self._solid_dirty = True
            return
        object.__setattr__(self, name, value)

    @property
    def solid_forcex(self) -> HostArray:
        return HostArray(self._forces[0, : self.n_objects, 0].cpu().numpy())

    @property
    def solid_forcey(self) -> HostArray:
        return HostArray(self._forces[0, : self.n_objects, 1].cpu().numpy())

    @property
    def torque(self) -> HostArray:
        return HostArray(self._forces[0, : self.n_objects, 5].cpu().numpy())

    @property
    def solid_line_transformed(self) -> HostArray:
        return HostArray(self._transformed()[0])

    # ------------------------------------------------------------ solids
    def configure_solid(self, solid_id: int, lines: Any, position: Any, angle: float = 0.0, scale: float = 0.3,
                        mass: float = 100000.0, mass_center: Optional[Any] = None) -> None:
        """Closed polygon body (lbm_core.py:172-282): vertices in body units,
        placed at `position` (lattice units) with angle and scale relative to nx."""
        if solid_id < 0 or solid_id >= self.n_objects:
            raise ValueError(f"solid_id {solid_id} out of range [0, {self.n_objects})")
        pts = as_numpy(lines, np.float32).reshape(-1, 2)
        n = len(pts)
        if n < 3:
            raise ValueError("a polygon needs at least 3 vertices")
        S = self._solid
        if n > self.max_segments_per_object:
            grown = np.zeros((S["solid_line"].shape[0], n, 2), np.float32)
            grown[:, : self.max_segments_per_object] = S["solid_line"].numpy()
            S["solid_line"] = HostArray(grown, on_change=self._mark_dirty)
            self.max_segments_per_object = n
        mc = np.zeros(2, np.float32) if mass_center is None else as_numpy(mass_center, np.float32).reshape(2)
        line = S["solid_line"].numpy()
        line[solid_id] = 0.0
        line[solid_id, :n] = pts
        S["solid_line"].assign(line)
        S["solid_line_num"][solid_id] = n
        S["solid_position"][solid_id] = as_numpy(position, np.float32).reshape(2)
        S["solid_mass_center"][solid_id] = mc
        S["solid_angle"][solid_id] = float(angle)
        S["solid_scale"][solid_id] = float(scale)
        S["mass"][solid_id] = float(mass)
        S["MoI"][solid_id] = float(mass) * (float(scale) * float(self.nx)) ** 2.0 / 8.0
        dist = np.sqrt(((pts.astype(np.float64) - mc) ** 2).sum(1)).max()
        S["solid_max_radius"][solid_id] = (dist + math.sqrt(2.0)) ** 2
        self._capsules.pop(solid_id, None)
        self._solid_dirty = True

    def add_capsule(self, solid_id: int, a: Sequence[float], b: Sequence[float], radius: float,
                    position: Optional[Sequence[float]] = None, angle: float = 0.0) -> None:
        """Capsule body (extension): segment a-b (body frame, lattice units)
        with the given radius, placed at position with angle."""
        if solid_id < 0 or solid_id >= self.n_objects:
            raise ValueError(f"solid_id {solid_id} out of range [0, {self.n_objects})")
        if not radius > 0:
            raise ValueError("capsule radius must be > 0")
        self._capsules[solid_id] = (np.asarray(a, np.float64), np.asarray(b, np.float64), float(radius))
        S = self._solid
        S["solid_line_num"][solid_id] = 0
        if position is not None:
            S["solid_position"][solid_id] = as_numpy(position, np.float32).reshape(2)
        S["solid_angle"][solid_id] = float(angle)
        self._solid_dirty = True

    def _transformed(self):
        """Polygon vertices in lattice coordinates, FP32, in the operation order
        of precompute_transformed_segments (lbm_func.py:509-542)."""
        S = self._solid
        n = S["solid_line"].shape[0]
        line, num = S["solid_line"].numpy(), S["solid_line_num"].numpy()
        out = np.zeros_like(line, dtype=np.float32)
        f32 = np.float32
        for k in range(n):
            if num[k] == 0:
                continue
            scale = f32(S["solid_scale"][k]) * f32(self.nx)
            ca, sa = f32(np.cos(f32(S["solid_angle"][k]))), f32(np.sin(f32(S["solid_angle"][k])))
            pp = (line[k, : num[k]] - S["solid_mass_center"][k].astype(f32)) * scale
            pos = S["solid_position"][k].astype(f32)
            out[k, : num[k], 0] = pp[:, 0] * ca - pp[:, 1] * sa + pos[0]
            out[k, : num[k], 1] = pp[:, 0] * sa + pp[:, 1] * ca + pos[1]
        return out, num

    def _poses(self) -> np.ndarray:
        S = self._solid
        n = max(self.n_objects, 1)
        p = np.zeros((n, 4), np.float32)
        p[:, 0:2] = S["solid_position"].numpy()[:n]
        ang = S["solid_angle"].numpy()[:n].astype(np.float64)
        p[:, 2], p[:, 3] = np.cos(ang), np.sin(ang)
        return p

    def _prepare(self, stream=None) -> None:
        """Rebuild mask and pose table from the solid tables when they changed;
        the previous step's pose becomes the 'last' pose (u_p = frame difference)."""
        L = self.lattice
        if self.n_objects == 0:
            return
        if not self._solid_dirty and self._at_rest:
            return  # the uploaded table (pose | same pose) and the mask still hold
        now = self._poses()
        last = now if self._last_pose is None else self._last_pose
        table = np.concatenate([now, last], axis=1)[: self.n_objects].astype(np.float32)
        if self._uploaded is None or not np.array_equal(table, self._uploaded):  # static bodies: upload once
            L.set_poses(table, stream=stream)
            self._uploaded = table
        self._at_rest = bool(np.array_equal(now, last))
        self._last_pose = now
        if self._uses_cut_cells():
            self._prepare_segments()
        if not self._solid_dirty:
            return
        verts, num = self._transformed()
        polys = [k for k in range(self.n_objects) if num[k] >= 3]
        caps = []
        for k, (a, b, r) in self._capsules.items():
            c, s = now[k, 2], now[k, 3]
            R = np.array([[c, -s], [s, c]], np.float64)
            pa = R @ a[:2] + now[k, :2]
            pb = R @ b[:2] + now[k, :2]
            caps.append([pa[0], pa[1], 0.0, pb[0], pb[1], 0.0, r, k])
        if polys and caps:
            raise ValueError("a world's bodies are either polygons (configure_solid) or capsules (add_capsule)")
        if polys:
            vv = np.zeros((self.n_objects, max(int(num.max()), 3), 2), np.float32)
            cnt = np.zeros(self.n_objects, np.int32)
            for k in polys:
                vv[k, : num[k]] = verts[k, : num[k]]
                cnt[k] = num[k]
            L.voxelize_polygons(vv, cnt, stream=stream)
        elif caps:
            L.voxelize(np.array(caps, np.float32), stream=stream)
        self._solid_dirty = False

    def _uses_cut_cells(self) -> bool:
        """Polygon bodies stepped with the reference's segment-link cut cells."""
        return self.cut_cell and self.n_objects > 0 and not self._capsules

    def _prepare_segments(self) -> None:
        """Device tables of the exact cut-cell step: this step's transformed
        vertices and the previous step's, as precompute_transformed_segments
        keeps them (lbm_func.py:509-542: a vertex whose stored position is
        exactly (0, 0) counts as new, so its last position is the new one),
        and the squared bounding radii (solid_max_radius * scale) * scale in
        FP32 with scale = solid_scale * nx (lbm_func.py:569-574)."""
        f32 = np.float32
        verts, num = self._transformed()
        prev = self._seg_prev
        if prev is None or prev.shape != verts.shape:
            grown = np.zeros_like(verts)
            if prev is not None:
                k = min(prev.shape[1], verts.shape[1])
                grown[:, :k] = prev[:, :k]
            prev = grown
        last = prev.copy()
        for k in range(verts.shape[0]):
            n = int(num[k])
            fresh = (prev[k, :n, 0] == 0.0) & (prev[k, :n, 1] == 0.0)
            last[k, :n] = np.where(fresh[:, None], verts[k, :n], prev[k, :n])
            prev[k, :n] = verts[k, :n]
        self._seg_prev = prev
        S = self._solid
        scale = S["solid_scale"].numpy().astype(f32) * f32(self.nx)
        r2 = (S["solid_max_radius"].numpy().astype(f32) * scale) * scale
        n = self.n_objects
        host = (np.ascontiguousarray(verts[:n], f32), np.ascontiguousarray(last[:n], f32),
                np.ascontiguousarray(num[:n], np.int32), np.ascontiguousarray(r2[:n], f32))
        old = self._seg_host
        if old is not None and all(a.shape == b.shape and np.array_equal(a, b) for a, b in zip(host, old)):
            return
        dev = self.lattice.device
        self._seg_dev = tuple(torch.from_numpy(h.copy()).to(dev) for h in host)
        self._seg_host = host

    def step(self, stream=None) -> None:
        """One time step with the per-solid force of the state it reads."""
        self._prepare(stream)
        self._launch(stream)

    def _launch(self, stream=None) -> None:
        """The device work of a step whose tables _prepare has set up."""
        L = self.lattice
        fh = self._forces if self.n_objects > 0 else None
        if self._uses_cut_cells():
            v, vl, cnt, r2 = self._seg_dev
            L.step_polygons(1, v, vl, cnt, r2, force_history=fh, stream=stream)
            return
        L.step(1, N.HL_KERNEL_AUTO, force_history=fh, stream=stream)


class LBM_Solver:
    """Worlds of one size stepped together (envs/lbm/lbm_solver.py:21-141)."""

    def __init__(self, nx: int, ny: int, solid_num: int = 1, nworld: int = 1, device: Optional[Any] = None,
                 dtype: str = "f32", batched: bool = True, cut_cell: bool = True) -> None:
        if nworld < 1:
            raise ValueError("nworld must be >= 1")
        self._use_batch = bool(batched)
        self._streams = None  # side streams of the world-by-world route
        self._batch = None
        self._batch_keys = None
        self._batch_boxes = None
        self._batch_cur = None
        self._hist = None
        self.nx, self.ny = nx, ny
        self.solid_num = solid_num
        self.nworld = nworld
        self.captured = False
        self.captured_graph = None
        self.total_steps = 500000
        self.per_frame_t = 500
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.flows = [HomeFlow(dtype=dtype, device=self.device) for _ in range(nworld)]
        for flow in self.flows:
            flow.Initialize(nx, ny, n_objects=solid_num)
            flow.cut_cell = bool(cut_cell)
        self.flows_wp = self.flows
        self.mujoco_mappings = {}
        self.mujoco_origins_wp = None
        self.lbm_origins_wp = None
        self.scales_wp = None
        self.solid_ids_wp = None
        self.solid_id_to_index = {}

    def _batchable(self) -> bool:
        from .batch import MAX_SOLIDS
        return self._use_batch and len(self.flows) > 1 and all(
            f.lattice.dtype == "f32" and f.lattice.force is None and f.lattice.nsolid <= MAX_SOLIDS
            and not f._uses_cut_cells()
            for f in self.flows)

    def _polygon_worlds(self) -> bool:
        """Several worlds, all on the exact cut cells, of one dtype and at
        the same buffer: the multi-world polygon route."""
        return self._use_batch and len(self.flows) > 1 and all(f._uses_cut_cells() for f in self.flows) and len(
            {(f.lattice.dtype_code, f.lattice.cur) for f in self.flows}) == 1

    @staticmethod
    def _descriptor_key(L):
        """What a batch keeps of a world's descriptor, except the mask box
        (updated in place): buffers, viscosity, faces."""
        return (L.mask.data_ptr() if L.mask is not None else 0, L.force is None, L.nsolid, float(L.nu),
                tuple(L.bc_type), tuple(tuple(v) for v in L.bc_value), L.buf.data_ptr(), L.solids.data_ptr())

    def step(self) -> None:
        """Advance every world one step (asynchronous): with several FP32 worlds
        on the mask rule and no body-force field, one batched launch sequence
        for all of them (bitwise the same as stepping each world alone), else
        world by world, the worlds spread over a few concurrent streams.
        The reference captures a graph on the first call; resetting `captured`
        is harmless here."""
        if not self._batchable():
            if len(self.flows) == 1:
                self.flows[0].step()
            elif self._polygon_worlds():
                # every world on the exact cut cells: one library call, one
                # gather and one band launch per eight worlds
                for flow in self.flows:
                    flow._prepare()
                step_polygons_worlds([f.lattice for f in self.flows], 1, [f._seg_dev for f in self.flows],
                                     [f._forces for f in self.flows])
            else:
                # worlds are independent: their launches go round-robin to two
                # side streams (default priority, each with its own library
                # workspace), forked from and joined back into the current
                # stream; host-side set-up and uploads stay on the current one.
                # Two: with the current stream that stays within the four
                # hardware queues HIP gives a process by default (more streams
                # than queues measured slower)
                cur = torch.cuda.current_stream(self.device)
                for flow in self.flows:
                    flow._prepare()
                if self._streams is None:
                    self._streams = [torch.cuda.Stream(device=self.device) for _ in range(min(2, len(self.flows)))]
                fork = torch.cuda.Event()
                fork.record(cur)
                for st in self._streams:
                    st.wait_event(fork)
                for w, flow in enumerate(self.flows):
                    flow._launch(self._streams[w % len(self._streams)])
                for st in self._streams:
                    join = torch.cuda.Event()
                    join.record(st)
                    cur.wait_event(join)
            self.captured = True
            return
        from .batch import Batch
        for flow in self.flows:
            flow._prepare()
        lats = [f.lattice for f in self.flows]
        keys = [self._descriptor_key(L) for L in lats]
        if self._batch is None or keys != self._batch_keys:
            if self._batch is not None:
                self._batch.close()
            self._batch = Batch(lats)
            self._batch_keys = keys
            self._batch_boxes = [tuple(L.mask_box) for L in lats]
            self._hist = self._batch.history_buffer(1)
            # each world's force rows are read from the batch's history in place
            for i, f in enumerate(self.flows):
                if f.n_objects > 0:
                    f._forces = self._hist[:, i, : f._forces.shape[1]]
        else:
            if [L.cur for L in lats] != self._batch_cur:  # a world was stepped or reset on its own
                self._batch._push_state()
            for i, L in enumerate(lats):
                if tuple(L.mask_box) != self._batch_boxes[i]:  # bodies moved: new mask box
                    self._batch.set_kinematics(i, None, None, box=L.mask_box)
                    self._batch_boxes[i] = tuple(L.mask_box)
        self._batch.step(1, force_history=self._hist)
        self._batch_cur = [L.cur for L in lats]
        self.captured = True

    def create_solid_from_mujoco(self, solid_id, model, data, body_or_geom_name, lbm_position=None, lbm_scale=0.3,
                                 n_samples=20, is_body=True, **kw):
        """Project one body or geom of an MJCF model into every world
        (lbm_solver.py:143-208, extract_lbm_polygon_from_mujoco,
        envs/lbm/mujoco_to_lbm.py:343-395): its x-y outline, convex hull and
        centring, placed at lbm_position (the grid centre by default) with its
        rotation about z and lbm_scale. `model` is an MJCF file path or an
        openhomelbm_amd.mjcf.Model (the file's reference configuration: every
        joint at 0, as after mj_forward on qpos0); `data` is unused. A MuJoCo
        MjModel is not accepted (MuJoCo does not run on this platform)."""
        from . import mjcf

        if isinstance(model, (str, os.PathLike)):
            model = mjcf.load_model(os.fspath(model))
        if not isinstance(model, mjcf.Model):
            raise NotImplementedError("create_solid_from_mujoco takes an MJCF path or openhomelbm_amd.mjcf.Model; "
                                      "MuJoCo models need MuJoCo, which is not available on this platform")
        info = mjcf.polygon_from_model(model, body_or_geom_name, n_samples=n_samples, use_convex_hull=True,
                                       normalize=True, is_body=is_body)
        if lbm_position is None:
            lbm_position = (self.nx * 0.5, self.ny * 0.5)
        for flow in self.flows:
            flow.configure_solid(solid_id, info["vertices"], lbm_position, angle=info["angle"], scale=lbm_scale, **kw)
        self.mujoco_mappings[solid_id] = {"mujoco_origin": np.array(info["position"], np.float64).copy(),
                                          "lbm_origin": np.array(lbm_position, np.float32),
                                          "scale": lbm_scale * self.nx}
        return info

    def finalize_mappings(self, solid_ids: Sequence[int]) -> None:
        """Coordinate-mapping tables for the coupling (lbm_solver.py:211-240)."""
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

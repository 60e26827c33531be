"""ctypes front end of the C oracle (oracle/homelbm_oracle.c).

TEST INFRASTRUCTURE ONLY. Imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, as the checker. The product never imports it.

Arrays are NumPy ``(nfields, nz, ny, nx)`` (D3Q27) or ``(nfields, ny, nx)``
(D2Q9), C-contiguous, so the field stride equals the cell count.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)
    return _LIB_PATH


class _Env(C.Structure):
    _fields_ = [
        ("lattice", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("stride", C.c_int64),
        ("mask", C.c_void_p),
        ("force", C.c_void_p),
        ("solids", C.c_void_p),
        ("nsolid", C.c_int32),
        ("nu", C.c_float),
        ("bc_type", C.c_int32 * 6),
        ("bc_value", (C.c_float * 3) * 6),
    ]


class _Polys(C.Structure):
    _fields_ = [
        ("verts", C.c_void_p),
        ("verts_last", C.c_void_p),
        ("counts", C.c_void_p),
        ("maxv", C.c_int32),
        ("radius2", C.c_void_p),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        for name in ("hlo_step_f32", "hlo_step_f64", "hlo_step_omp_f32",
                     "hlo_forces_f32", "hlo_forces_f64"):
            getattr(L, name).argtypes = [C.POINTER(_Env), P, P]
            getattr(L, name).restype = None
        L.hlo_init_uniform_f32.argtypes = [C.POINTER(_Env), C.c_float, C.c_float, C.c_float, P]
        L.hlo_init_uniform_f64.argtypes = [C.POINTER(_Env), C.c_double, C.c_double, C.c_double, P]
        L.hlo_recon_f32.argtypes = [C.c_int, P, C.c_int]
        L.hlo_recon_f32.restype = C.c_float
        L.hlo_recon_f64.argtypes = [C.c_int, P, C.c_int]
        L.hlo_recon_f64.restype = C.c_double
        L.hlo_voxelize.argtypes = [C.POINTER(_Env), P, C.c_int32, P, P]
        L.hlo_voxelize_polygons.argtypes = [C.POINTER(_Env), P, P, C.c_int32, C.c_int32, P, P]
        for name in ("hlo_step_polygons_f32", "hlo_step_polygons_f64"):
            getattr(L, name).argtypes = [C.POINTER(_Env), C.POINTER(_Polys), P, P, P]
            getattr(L, name).restype = None
        L.hlo_capsule_cut_f32.argtypes = [C.POINTER(_Env), P, C.c_int32, C.c_float, C.c_float, C.c_float,
                                          C.c_int32, C.POINTER(C.c_float)]
        L.hlo_capsule_cut_f32.restype = C.c_int32
        L.hlo_capsule_cut_f64.argtypes = [C.POINTER(_Env), P, C.c_int32, C.c_double, C.c_double, C.c_double,
                                          C.c_int32, C.POINTER(C.c_double)]
        L.hlo_capsule_cut_f64.restype = C.c_int32
        for name in ("hlo_step_capsules_f32", "hlo_step_capsules_f64"):
            getattr(L, name).argtypes = [C.POINTER(_Env), P, C.c_int32, P, P, P]
            getattr(L, name).restype = None
        L.hlo_step_capsules_omp_f32.argtypes = [C.POINTER(_Env), P, C.c_int32, P, P]
        L.hlo_step_capsules_omp_f32.restype = None
        L.hlo_mass_f32.argtypes = [C.POINTER(_Env), P]
        L.hlo_mass_f32.restype = C.c_double
        _lib = L
    return _lib


@dataclass
class Env:
    """Description of one lattice for the oracle.

    bc_type / bc_value use the reference's face order: D2Q9 (left, top,
    right, bottom) as in envs/lbm/lbm_core.py:136-141, D3Q27 (left, right,
    top, bottom, front, back) as in envs/lbm3d/lbm_core_3d.py:157-165.
    """

    lattice: int
    nx: int
    ny: int
    nz: int = 1
    nu: float = 0.01
    bc_type: Sequence[int] = (1, 1, 1, 1, 1, 1)
    bc_value: Sequence[Sequence[float]] = ((0.0, 0.0, 0.0),) * 6
    mask: Optional[np.ndarray] = None
    force: Optional[np.ndarray] = None
    solids: Optional[np.ndarray] = None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def nfields(self) -> int:
        return 6 if self.lattice == 9 else 10

    @property
    def shape(self):
        if self.lattice == 9:
            return (self.ny, self.nx)
        return (self.nz, self.ny, self.nx)

    @property
    def ncell(self) -> int:
        return self.nx * self.ny * self.nz

    def c_env(self) -> _Env:
        e = _Env()
        e.lattice = self.lattice
        e.nx, e.ny, e.nz = self.nx, self.ny, (self.nz if self.lattice == 27 else 1)
        e.stride = self.ncell
        self._keep = []
        if self.mask is not None:
            m = np.ascontiguousarray(self.mask, dtype=np.uint8)
            assert m.size == self.ncell
            self._keep.append(m)
            e.mask = m.ctypes.data
        if self.force is not None:
            f = np.ascontiguousarray(self.force, dtype=np.float32)
            assert f.size == (2 if self.lattice == 9 else 3) * self.ncell
            self._keep.append(f)
            e.force = f.ctypes.data
        if self.solids is not None:
            s = np.ascontiguousarray(self.solids, dtype=np.float32)
            self._keep.append(s)
            e.solids = s.ctypes.data
            e.nsolid = s.shape[0]
        e.nu = self.nu
        for k in range(6):
            e.bc_type[k] = int(self.bc_type[k]) if k < len(self.bc_type) else 1
            v = self.bc_value[k] if k < len(self.bc_value) else (0.0, 0.0, 0.0)
            for a in range(3):
                e.bc_value[k][a] = float(v[a]) if a < len(v) else 0.0
        return e


def _check(env: Env, a: np.ndarray, dtype) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=dtype)
    assert a.shape == (env.nfields,) + env.shape, (a.shape, env.shape)
    return a


def step(env: Env, src: np.ndarray, precision: str = "f32", omp: bool = False) -> np.ndarray:
    dtype = np.float32 if precision == "f32" else np.float64
    src = _check(env, src, dtype)
    dst = np.empty_like(src)
    e = env.c_env()
    if precision == "f32":
        fn = lib().hlo_step_omp_f32 if omp else lib().hlo_step_f32
    else:
        fn = lib().hlo_step_f64
    fn(C.byref(e), src.ctypes.data, dst.ctypes.data)
    return dst


def run(env: Env, src: np.ndarray, nsteps: int, precision: str = "f32", omp: bool = False,
        forces: bool = False):
    """Advance nsteps; optionally return the per-step force history."""
    cur = src
    hist = []
    for _ in range(nsteps):
        if forces:
            hist.append(compute_forces(env, cur, precision))
        cur = step(env, cur, precision, omp)
    if forces:
        return cur, np.array(hist)
    return cur


def compute_forces(env: Env, src: np.ndarray, precision: str = "f32") -> np.ndarray:
    dtype = np.float32 if precision == "f32" else np.float64
    src = _check(env, src, dtype)
    e = env.c_env()
    out = np.zeros((max(e.nsolid, 1), 6), dtype=np.float64)
    fn = lib().hlo_forces_f32 if precision == "f32" else lib().hlo_forces_f64
    fn(C.byref(e), src.ctypes.data, out.ctypes.data)
    return out[: e.nsolid]


def step_polygons(env: Env, src: np.ndarray, verts: np.ndarray, verts_last: np.ndarray, counts: Sequence[int],
                  radius2: Sequence[float], precision: str = "f32"):
    """One D2Q9 step with the reference's segment-link cut cells
    (get_cutcell, lbm_func.py:545-607): verts / verts_last [nsolid, maxv, 2]
    FP32 transformed vertices of this and the previous step, counts, radius2
    the squared bounding radii. Returns (state, forces [nsolid, 6]) with the
    forces of the state read. env.mask is ignored."""
    dtype = np.float32 if precision == "f32" else np.float64
    src = _check(env, src, dtype)
    dst = np.empty_like(src)
    e = env.c_env()
    e.mask = None
    v = np.ascontiguousarray(verts, np.float32)
    vl = np.ascontiguousarray(verts_last, np.float32)
    assert v.shape == vl.shape and v.ndim == 3 and v.shape[0] == e.nsolid and v.shape[2] == 2
    cnt = np.ascontiguousarray(counts, np.int32)
    r2 = np.ascontiguousarray(radius2, np.float32)
    pol = _Polys(v.ctypes.data, vl.ctypes.data, cnt.ctypes.data, v.shape[1], r2.ctypes.data)
    out = np.zeros((max(e.nsolid, 1), 6), dtype=np.float64)
    fn = lib().hlo_step_polygons_f32 if precision == "f32" else lib().hlo_step_polygons_f64
    fn(C.byref(e), C.byref(pol), src.ctypes.data, dst.ctypes.data, out.ctypes.data)
    return dst, out[: e.nsolid]


def _caps(capsules) -> np.ndarray:
    return np.ascontiguousarray(capsules, dtype=np.float32).reshape(-1, 8)


def capsule_cut(env: Env, capsules: np.ndarray, cell: Sequence[float], i: int, precision: str = "f32"):
    """(solid, t) of the crossing of link i from cell towards cell - c_i
    nearest to the cell over every capsule surface (solid -1: none)."""
    caps = _caps(capsules)
    e = env.c_env()
    if precision == "f32":
        t = C.c_float(0)
        sid = lib().hlo_capsule_cut_f32(C.byref(e), caps.ctypes.data, caps.shape[0], *[float(v) for v in cell],
                                        int(i), C.byref(t))
    else:
        t = C.c_double(0)
        sid = lib().hlo_capsule_cut_f64(C.byref(e), caps.ctypes.data, caps.shape[0], *[float(v) for v in cell],
                                        int(i), C.byref(t))
    return int(sid), float(t.value)


def step_capsules(env: Env, src: np.ndarray, capsules: np.ndarray, precision: str = "f32", omp: bool = False):
    """One D3Q27 step whose cut links are the capsules' crossings (the ray
    query of get_cutcell_multi_3d with closed-form capsule surfaces) and whose
    wall point is the crossing (lbm_func_3d.py:424-511). Returns (state,
    forces [nsolid, 6] of the state read; None with omp). env.mask is
    ignored."""
    dtype = np.float32 if precision == "f32" else np.float64
    src = _check(env, src, dtype)
    dst = np.empty_like(src)
    caps = _caps(capsules)
    e = env.c_env()
    if omp:
        assert precision == "f32"
        lib().hlo_step_capsules_omp_f32(C.byref(e), caps.ctypes.data, caps.shape[0], src.ctypes.data,
                                        dst.ctypes.data)
        return dst, None
    out = np.zeros((max(e.nsolid, 1), 6), dtype=np.float64)
    fn = lib().hlo_step_capsules_f32 if precision == "f32" else lib().hlo_step_capsules_f64
    fn(C.byref(e), caps.ctypes.data, caps.shape[0], src.ctypes.data, dst.ctypes.data, out.ctypes.data)
    return dst, out[: e.nsolid]


def init_uniform(env: Env, u: Sequence[float], precision: str = "f32") -> np.ndarray:
    dtype = np.float32 if precision == "f32" else np.float64
    dst = np.empty((env.nfields,) + env.shape, dtype=dtype)
    e = env.c_env()
    u = list(u) + [0.0] * (3 - len(u))
    fn = lib().hlo_init_uniform_f32 if precision == "f32" else lib().hlo_init_uniform_f64
    fn(C.byref(e), u[0], u[1], u[2], dst.ctypes.data)
    return dst


def recon(lattice: int, m: Sequence[float], i: int, precision: str = "f64") -> float:
    dtype = np.float32 if precision == "f32" else np.float64
    a = np.ascontiguousarray(m, dtype=dtype)
    fn = lib().hlo_recon_f32 if precision == "f32" else lib().hlo_recon_f64
    return fn(lattice, a.ctypes.data, i)


def voxelize(env: Env, capsules: np.ndarray, box: Optional[Sequence[int]] = None,
             mask: Optional[np.ndarray] = None) -> np.ndarray:
    caps = np.ascontiguousarray(capsules, dtype=np.float32).reshape(-1, 8)
    if box is None:
        box = (0, env.nx, 0, env.ny, 0, env.nz if env.lattice == 27 else 1)
    b = np.ascontiguousarray(box, dtype=np.int32)
    if mask is None:
        mask = np.zeros(env.shape, dtype=np.uint8)
    assert mask.dtype == np.uint8 and mask.flags.c_contiguous and mask.size == env.ncell
    e = env.c_env()
    lib().hlo_voxelize(C.byref(e), caps.ctypes.data, caps.shape[0], b.ctypes.data, mask.ctypes.data)
    return mask


def voxelize_polygons(env: Env, verts: np.ndarray, counts: Sequence[int], box: Optional[Sequence[int]] = None,
                      mask: Optional[np.ndarray] = None) -> np.ndarray:
    """verts: [npoly, maxv, 2] FP32 lattice coordinates; counts: vertices per polygon."""
    v = np.ascontiguousarray(verts, dtype=np.float32)
    c = np.ascontiguousarray(counts, dtype=np.int32)
    if box is None:
        box = (0, env.nx, 0, env.ny, 0, 1)
    b = np.ascontiguousarray(box, dtype=np.int32)
    if mask is None:
        mask = np.zeros(env.shape, dtype=np.uint8)
    e = env.c_env()
    lib().hlo_voxelize_polygons(C.byref(e), v.ctypes.data, c.ctypes.data, v.shape[0], v.shape[1], b.ctypes.data,
                                mask.ctypes.data)
    return mask


def mass(env: Env, src: np.ndarray) -> float:
    src = _check(env, src, np.float32)
    e = env.c_env()
    return lib().hlo_mass_f32(C.byref(e), src.ctypes.data)

"""CPU restatement of the render / diagnostic kernels (test infrastructure
only: imported by tests, never by the package).

Follows, in FP32 and in the same operation order:
  get_u_img                        envs/lbm/lbm_func.py:163-166
  get_velocity_rgb                 envs/lbm/lbm_func.py:169-193
  get_vorticity_img                envs/lbm/lbm_func.py:196-233
  get_vorticity_with_solid_img     envs/lbm/lbm_func.py:236-319 (bodies: voxel mask)
  get_solid_boundary_img           envs/lbm/lbm_func.py:804-864 (bodies: voxel mask)
  get_u_img_xy/yz_3d               envs/lbm3d/lbm_func_3d.py:753-776
  get_vorticity_xy/yz_3d(_with_solid) envs/lbm3d/lbm_func_3d.py:779-988
  get_u_projection_*_3d, get_vorticity_projection_*_3d
                                   envs/lbm3d/lbm_func_3d.py:990-1313 (rays: voxel mask)
The squares u**2.0 are products. The reference marks bodies with polygon /
mesh distance queries; here (as in the GPU kernels) a cell is solid where the
voxel mask is set or a masked cell lies within distance `thickness`, and a
projected ray is solid where it meets the mask: parity for the body overlay
is against this definition, not the reference's geometry.

Arrays: velocity u[c] with shape [ny, nx] (2D) or [nz, ny, nx] (3D), x
fastest; images are returned in the same (row = second axis) layout as the
GPU kernels: plane xy -> [ny, nx], plane yz -> [nz, ny], projection along x
-> [nz, ny], along y -> [nz, nx], along z -> [ny, nx].
"""
from __future__ import annotations

import numpy as np

F = np.float32
SOLID = F(1000.0)


def grid_length(nx: int) -> np.float32:
    """flow.grid_length = 1.0 / nx stored in an FP32 struct field (lbm_core.py:105)."""
    return F(1.0 / nx)


def deriv(v: np.ndarray, axis: int, gl: np.float32) -> np.ndarray:
    """d v / d axis: central inside, forward at 0, backward at n-1
    (lbm_func.py:208-230). axis counts x=0, y=1, z=2."""
    v = v.astype(F)
    ax = v.ndim - 1 - axis
    n = v.shape[ax]
    out = np.empty_like(v)

    def sl(i):
        s = [slice(None)] * v.ndim
        s[ax] = i
        return tuple(s)

    two_gl = F(2.0) * gl
    out[sl(slice(1, n - 1))] = (v[sl(slice(2, n))] - v[sl(slice(0, n - 2))]) / two_gl
    out[sl(0)] = (v[sl(1)] - v[sl(0)]) / gl
    out[sl(n - 1)] = (v[sl(n - 1)] - v[sl(n - 2)]) / gl
    return out


def curl(u: np.ndarray, normal: int, gl: np.float32) -> np.ndarray:
    """Vorticity about `normal`: 2 -> dv/dx - du/dy (lbm_func.py:233;
    lbm_func_3d.py:808), 0 -> dw/dy - dv/dz (lbm_func_3d.py:840),
    1 -> du/dz - dw/dx (lbm_func_3d.py:1289-1307)."""
    if normal == 2:
        return deriv(u[1], 0, gl) - deriv(u[0], 1, gl)
    if normal == 0:
        return deriv(u[2], 1, gl) - deriv(u[1], 2, gl)
    return deriv(u[0], 2, gl) - deriv(u[2], 0, gl)


def speed(u: np.ndarray) -> np.ndarray:
    u = u.astype(F)
    s = u[0] * u[0] + u[1] * u[1]
    if len(u) == 3:
        s = s + u[2] * u[2]
    return np.sqrt(s)


def velocity_rgb(u: np.ndarray, vmax: float) -> np.ndarray:
    """(R, G, B) of lbm_func.py:169-193, channel last."""
    u = u.astype(F)
    vm = F(vmax)
    mag = np.sqrt(u[0] * u[0] + u[1] * u[1])
    r = np.clip(mag / vm, F(0), F(1))
    g = np.clip((u[0] + vm) / (F(2.0) * vm), F(0), F(1))
    b = np.clip((u[1] + vm) / (F(2.0) * vm), F(0), F(1))
    return np.stack([r, g, b], axis=-1)


def solid_near(mask: np.ndarray, thickness: float) -> np.ndarray:
    """Masked, or a masked cell within distance t (cell centres; 3D offsets
    for 3D masks)."""
    m = mask != 0
    out = m.copy()
    t = F(thickness)
    if not t > 0:
        return out
    r = int(t)
    offs = range(-r, r + 1)
    dz_range = offs if m.ndim == 3 else [0]
    for dz in dz_range:
        for dy in offs:
            for dx in offs:
                if F(dx * dx + dy * dy + dz * dz) > t * t:
                    continue
                sh = np.zeros_like(m)
                src = m
                if m.ndim == 2:
                    sh[max(0, -dy):m.shape[0] - max(0, dy), max(0, -dx):m.shape[1] - max(0, dx)] = \
                        src[max(0, dy):m.shape[0] - max(0, -dy), max(0, dx):m.shape[1] - max(0, -dx)]
                else:
                    sh[max(0, -dz):m.shape[0] - max(0, dz), max(0, -dy):m.shape[1] - max(0, dy),
                       max(0, -dx):m.shape[2] - max(0, dx)] = \
                        src[max(0, dz):m.shape[0] - max(0, -dz), max(0, dy):m.shape[1] - max(0, -dy),
                            max(0, dx):m.shape[2] - max(0, -dx)]
                out |= sh
    return out


def face_cells(shape) -> np.ndarray:
    """ML_WALL flags of InitBoundary3D (lbm_func_3d.py:13-19)."""
    f = np.zeros(shape, bool)
    f[0], f[-1] = True, True
    f[:, 0], f[:, -1] = True, True
    if len(shape) == 3:
        f[:, :, 0], f[:, :, -1] = True, True
    return f


# ----------------------------------------------------------------- planes
def plane(u: np.ndarray, kind: str, normal: int = 2, index: int = 0, mask=None, thickness: float = 0.0,
          vmax: float = 1.0) -> np.ndarray:
    """One image as the GPU kernel lays it out. kind: speed, vorticity,
    vorticity_solid, rgb, solid."""
    gl = grid_length(u.shape[-1])
    three = u.ndim == 4

    def take(a):
        if not three:
            return a
        return a[index] if normal == 2 else a[:, :, index]

    if kind == "speed":
        return take(speed(u))
    if kind == "rgb":
        return take(velocity_rgb(u, vmax))
    if kind in ("vorticity", "vorticity_solid"):
        img = take(curl(u, normal if three else 2, gl))
        if kind == "vorticity_solid" and mask is not None:
            img = np.where(take(solid_near(mask, thickness)), SOLID, img)
        return img
    if kind == "solid":
        s = np.zeros(u.shape[1:], bool) if mask is None else solid_near(mask, thickness)
        return take(s).astype(F)
    raise ValueError(kind)


# ----------------------------------------------------------------- projections
def projection(u: np.ndarray, kind: str, axis: int, mask=None) -> np.ndarray:
    """One ray per pixel along axis (x=0, y=1, z=2) of a 3D field.
    kind: first_speed (from the high-index end, first |u| > 1e-8), max_speed,
    max_vorticity (signed value of the largest |curl| about the axis over the
    interior, first one wins ties); face cells skipped; a ray through the
    mask reads 1000."""
    u = u.astype(F)
    gl = grid_length(u.shape[-1])
    wall = face_cells(u.shape[1:])
    ax = 2 - axis  # numpy axis of the ray
    sp = speed(u)
    q = curl(u, axis, gl) if kind == "max_vorticity" else None
    val = np.moveaxis(sp if q is None else q, ax, -1)
    wl = np.moveaxis(wall, ax, -1)
    n = val.shape[-1]
    out = np.zeros(val.shape[:-1], F)
    if kind == "first_speed":
        done = np.zeros(out.shape, bool)
        for k in range(n - 1, -1, -1):
            take = ~done & ~wl[..., k] & (val[..., k] > F(1e-8))
            out[take] = val[..., k][take]
            done |= take
    elif kind == "max_speed":
        for k in range(n - 1, -1, -1):
            v = np.where(wl[..., k], F(0), val[..., k])
            out = np.where(v > out, v, out)
    elif kind == "max_vorticity":
        for k in range(1, n - 1):
            v = val[..., k]
            take = ~wl[..., k] & (np.abs(v) > np.abs(out))
            out = np.where(take, v, out)
    else:
        raise ValueError(kind)
    if mask is not None:
        hit = np.moveaxis(mask != 0, ax, -1).any(-1)
        out = np.where(hit, SOLID, out)
    # image rows: along x -> [z][y], along y -> [z][x], along z -> [y][x]
    return out

"""Render and diagnostic kernels of the reference, by name, over the HIP
library (envs/lbm/lbm_func.py:163-319, 804-864; envs/lbm3d/lbm_func_3d.py:753-1313).

Each function takes the flow (HomeFlow / HomeFlow3D) and the kernel's scalar
inputs, and writes the same image buffer the reference kernel writes:
``flow.u_img`` / ``flow.u_img_rgb`` (2D), ``flow.u_img_xy``, ``flow.u_img_xz``
and ``flow.u_img_xz_front`` (3D), read back with ``.numpy()`` in the
reference's index order ([x, y], [y, z], [x, z]). A caller's

    wp.launch(get_vorticity_img, dim=(flow.nx, flow.ny), inputs=[flow])

becomes ``launch(get_vorticity_img, dim=(flow.nx, flow.ny), inputs=[flow])``
with ``launch`` from this module.

Bodies are the flow's voxel mask (solid where masked, or with thickness t
within distance t of a masked cell), where the reference queries polygon or
mesh distances; the fluid values are the reference's formulas.
"""
from __future__ import annotations

from typing import Any, Callable, Optional, Sequence

from . import _native as N


def launch(kernel: Callable, dim: Any = None, inputs: Sequence[Any] = (), device: Optional[Any] = None,
           **_ignored) -> None:
    """wp.launch-shaped call: the grid is implied by the flow."""
    kernel(*inputs)


def _set(flow, name: str, img) -> None:
    flow._images[name] = img


# ----------------------------------------------------------------- 2D
def get_u_img(flow) -> None:
    """|u| into u_img (lbm_func.py:163-166)."""
    _set(flow, "u_img", flow.lattice.render_plane(N.HL_RENDER_SPEED).T)


def get_velocity_rgb(flow, max_velocity: float) -> None:
    """(|u|, ux, uy) scaled to [0, 1] into u_img_rgb (lbm_func.py:169-193)."""
    img = flow.lattice.render_plane(N.HL_RENDER_RGB, max_velocity=max_velocity)
    _set(flow, "u_img_rgb", img.permute(1, 0, 2))


def get_vorticity_img(flow) -> None:
    """dv/dx - du/dy into u_img (lbm_func.py:196-233)."""
    _set(flow, "u_img", flow.lattice.render_plane(N.HL_RENDER_VORTICITY).T)


def get_vorticity_with_solid_img(flow, thickness: float) -> None:
    """Vorticity, 1000 on bodies (lbm_func.py:236-319)."""
    _set(flow, "u_img", flow.lattice.render_plane(N.HL_RENDER_VORTICITY_SOLID, thickness=thickness).T)


def get_solid_boundary_img(flow, thickness: float) -> None:
    """1 on bodies, 0 elsewhere, into u_img (lbm_func.py:804-864)."""
    _set(flow, "u_img", flow.lattice.render_plane(N.HL_RENDER_SOLID, thickness=thickness).T)


# ----------------------------------------------------------------- 3D slices
def get_u_img_xy_3d(flow, z_slice: int) -> None:
    """|u| on the plane z = z_slice into u_img_xy (lbm_func_3d.py:753-761)."""
    _set(flow, "u_img_xy", flow.lattice.render_plane(N.HL_RENDER_SPEED, 2, z_slice).T)


def get_u_img_yz_3d(flow, x_slice: int) -> None:
    """|u| on the plane x = x_slice into u_img_xz, read as [y, z] (lbm_func_3d.py:764-776)."""
    _set(flow, "u_img_xz", flow.lattice.render_plane(N.HL_RENDER_SPEED, 0, x_slice).T)


def get_vorticity_xy_3d(flow, z_slice: int) -> None:
    """curl_z on z = z_slice into u_img_xy (lbm_func_3d.py:779-808)."""
    _set(flow, "u_img_xy", flow.lattice.render_plane(N.HL_RENDER_VORTICITY, 2, z_slice).T)


def get_vorticity_yz_3d(flow, x_slice: int) -> None:
    """curl_x on x = x_slice into u_img_xz, [y, z] (lbm_func_3d.py:811-840)."""
    _set(flow, "u_img_xz", flow.lattice.render_plane(N.HL_RENDER_VORTICITY, 0, x_slice).T)


def get_vorticity_xy_with_solid_3d(flow, z_slice: int, thickness: float) -> None:
    """curl_z with bodies at 1000 (lbm_func_3d.py:843-911)."""
    img = flow.lattice.render_plane(N.HL_RENDER_VORTICITY_SOLID, 2, z_slice, thickness=thickness)
    _set(flow, "u_img_xy", img.T)


def get_vorticity_yz_with_solid_3d(flow, x_slice: int, thickness: float) -> None:
    """curl_x with bodies at 1000 (lbm_func_3d.py:914-988)."""
    img = flow.lattice.render_plane(N.HL_RENDER_VORTICITY_SOLID, 0, x_slice, thickness=thickness)
    _set(flow, "u_img_xz", img.T)


# ----------------------------------------------------------------- 3D projections
def get_u_projection_topdown_3d(flow) -> None:
    """First |u| > 1e-8 from the top along -z into u_img_xy (lbm_func_3d.py:990-1030)."""
    _set(flow, "u_img_xy", flow.lattice.render_projection(N.HL_PROJECT_FIRST_SPEED, 2).T)


def get_u_projection_max_topdown_3d(flow) -> None:
    """Largest |u| along z into u_img_xy (lbm_func_3d.py:1033-1069)."""
    _set(flow, "u_img_xy", flow.lattice.render_projection(N.HL_PROJECT_MAX_SPEED, 2).T)


def get_vorticity_projection_topdown_3d(flow) -> None:
    """Signed curl_z of largest magnitude along z into u_img_xy (lbm_func_3d.py:1072-1125)."""
    _set(flow, "u_img_xy", flow.lattice.render_projection(N.HL_PROJECT_MAX_VORTICITY, 2).T)


def get_u_projection_side_3d(flow) -> None:
    """Largest |u| along x into u_img_xz, [y, z] (lbm_func_3d.py:1128-1163)."""
    _set(flow, "u_img_xz", flow.lattice.render_projection(N.HL_PROJECT_MAX_SPEED, 0).T)


def get_u_projection_front_3d(flow) -> None:
    """Largest |u| along y into u_img_xz_front, [x, z] (lbm_func_3d.py:1166-1201)."""
    _set(flow, "u_img_xz_front", flow.lattice.render_projection(N.HL_PROJECT_MAX_SPEED, 1).T)


def get_vorticity_projection_side_3d(flow) -> None:
    """Signed curl_x of largest magnitude along x into u_img_xz (lbm_func_3d.py:1204-1257)."""
    _set(flow, "u_img_xz", flow.lattice.render_projection(N.HL_PROJECT_MAX_VORTICITY, 0).T)


def get_vorticity_projection_front_3d(flow) -> None:
    """Signed curl_y of largest magnitude along y into u_img_xz_front (lbm_func_3d.py:1260-1313)."""
    _set(flow, "u_img_xz_front", flow.lattice.render_projection(N.HL_PROJECT_MAX_VORTICITY, 1).T)

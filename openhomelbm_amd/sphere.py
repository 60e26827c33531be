"""Flow past the sphere workload (D3Q27, 256x128x128, D = 24 at (64, 64, 65),
U = 0.05, Re = 200): force histories under either 3D boundary rule and
their comparison, the 3D counterpart of karman.compare_boundaries.

The two rules:
  "capsules": the reference's own rule, a link cut where the ray from the
      cell along -c_i meets the sphere at t in [0, 1], u_p and the lever arm
      at the crossing x - c_i t (get_cutcell_multi_3d and its cut branch,
      envs/lbm3d/lbm_func_3d.py:254-316, 424-511, with the sphere in closed
      form instead of a triangle mesh);
  "mask": the integer voxel mask, links cut where the mask's occupancy
      changes, the wall point at the link midpoint.

No reference counterpart: the reference records per-step forces inside its
environments but reports no drag coefficients.
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np


def run(steps: int, boundary: str = "capsules", dtype: str = "f32", w=None, log=None) -> np.ndarray:
    """FP64 force/torque history [steps, 6] of the sphere under `boundary`."""
    import torch

    from . import inputs
    from .lattice import capsule_box

    if boundary not in ("capsules", "mask"):
        raise ValueError("boundary must be 'capsules' or 'mask'")
    w = w or inputs.d3q27_sphere()
    lat = inputs.build_lattice(w, dtype)
    hist = torch.zeros((steps, 1, 6), dtype=torch.float64, device=lat.device)
    caps = torch.from_numpy(w.capsules).to(lat.device)
    box = capsule_box(w.capsules, (w.nx, w.ny, w.nz))
    chunk = 2000
    for k0 in range(0, steps, chunk):
        k1 = min(steps, k0 + chunk)
        if boundary == "capsules":
            lat.step_capsules(k1 - k0, caps, box, force_history=hist[k0:k1])
        else:
            lat.step(k1 - k0, force_history=hist[k0:k1])
        if log:
            log(f"[sphere] {boundary}: step {k1}/{steps}")
    h = hist[:, 0].cpu().numpy()
    if not np.isfinite(h).all():
        raise RuntimeError(f"sphere run ({boundary}) diverged")
    return h


def drag_coefficient(fx, u: float = 0.05, diameter: float = 24.0, rho: float = 1.0):
    """C_d = 2 F_x / (rho U^2 pi D^2 / 4)."""
    return 2.0 * np.asarray(fx) / (rho * u * u * np.pi * diameter * diameter / 4.0)


def compare_boundaries(steps: int = 4000, window: int = 1000, log=None) -> Dict:
    """Both rules from the same initial state: mean drag coefficient, mean
    side force and mean torque magnitude over the last `window` steps, and the
    mask rule's relative drag difference from the crossing rule."""
    out: Dict[str, dict] = {}
    for b in ("capsules", "mask"):
        h = run(steps, b, log=log)[-window:]
        cd = drag_coefficient(h[:, 0])
        out[b] = {"mean_drag": float(h[:, 0].mean()), "cd": float(cd.mean()),
                  "cd_drift_last_window": float(cd[-1] - cd[0]),
                  "mean_side_force": float(np.hypot(h[:, 1], h[:, 2]).mean()),
                  "mean_torque": float(np.linalg.norm(h[:, 3:6], axis=1).mean())}
    ref = out["capsules"]
    return {"workload": "d3q27_sphere_256x128x128", "steps": steps, "window": window, "re": 200,
            "capsules": ref, "mask": out["mask"],
            "mask_rel_diff": {"mean_drag": abs(out["mask"]["mean_drag"] - ref["mean_drag"]) / abs(ref["mean_drag"])}}

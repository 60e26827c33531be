"""Precision sweep: the same D3Q27 run with moments stored in FP64 (the GPU
oracle: the gather kernel in FP64), FP32, FP16 and BF16 (16-bit storage of
rho - 1, u - U (the inlet flow) and S, each scaled per field; FP16 rounded to
nearest, BF16 stochastically; collision in FP32).

Every reported error is against the FP64 run of the same inputs: relative L2
of the velocity field after the last step, and the relative error of the
mean drag (sum of the chain's Fx over its solids, averaged over the second
half of the run).
"""
from __future__ import annotations

from typing import Dict, Iterable, Optional

import numpy as np
import torch

from . import inputs
from .lattice import Lattice


def run(eel: "inputs.Eel3D", dtype: str, steps: int, seed: int = 4004, re: float = 500.0, log=None,
        boundary: str = "mask", history: bool = False):
    """Advance the eel workload `steps` steps in `dtype`; returns (final u as
    an FP64 device tensor [3, nz, ny, nx], drag history [steps] FP64, and
    with history=True the full force history [steps, nseg, 6]). boundary
    "mask": the capsules voxelized every step, the mask rule; "capsules":
    links cut where they cross a capsule surface, the wall point at the
    crossing (Lattice.step_capsules)."""
    if boundary not in ("mask", "capsules"):
        raise ValueError("boundary must be 'mask' or 'capsules'")
    nseg = eel.nseg
    L = Lattice(27, eel.nx, eel.ny, eel.nz, dtype=dtype, nsolid=nseg)
    L.nu = eel.nu(re)
    L.bc_type = [0, 1, 2, 2, 2, 2]
    L.bc_value = [[inputs.EEL_U, 0.0, 0.0]] + [[0.0, 0.0, 0.0]] * 5
    L.init_uniform((inputs.EEL_U, 0.0, 0.0))
    rng = np.random.Generator(np.random.PCG64(seed))
    noise = torch.from_numpy((1e-3 * (2.0 * rng.random((3, eel.nz, eel.ny, eel.nx)) - 1.0)).astype(np.float32))
    noise = noise.to(L.device)
    for b in (0, 1):
        f = L.fields(b)
        for k in range(3):
            f[1 + k].copy_(L.encode(1 + k, L.decode(1 + k, f[1 + k]) + noise[k]))
    caps, poses = eel.tables(0, steps)
    box = eel.box(0, steps)
    caps_d = torch.from_numpy(caps).to(L.device)
    poses_d = torch.from_numpy(poses).to(L.device)
    hist = torch.zeros((steps, nseg, 6), dtype=torch.float64, device=L.device)
    for t in range(steps):
        L.solids[:nseg].copy_(poses_d[t])
        if boundary == "capsules":
            L.step_capsules(1, caps_d[t], eel.box(t, 1), force_history=hist[t:t + 1])
        else:
            L.voxelize_device(caps_d[t], box)
            L.step(1, force_history=hist[t:t + 1])
        if log and (t + 1) % 1000 == 0:
            log(f"[precision] {dtype}: step {t + 1}/{steps}")
    raw = L.fields()
    u = torch.stack([L.decode(1 + k, raw[1 + k]).to(torch.float64) for k in range(3)])
    drag = hist[:, :, 0].sum(1)
    if not bool(torch.isfinite(u).all()):
        raise RuntimeError(f"{dtype} run diverged")
    if history:
        return u, drag, hist.cpu().numpy()
    return u, drag


def compare_boundaries(steps: int = 4000, eel=None, log=None) -> Dict:
    """The swimming eel in FP32 with capsule-crossing cut links and with the
    mask rule, from the same inputs: mean drag (sum of Fx over the chain),
    amplitude of the lateral force (sum of Fy) and of the yaw torque (sum of
    Tz), over the second half, and the mask rule's relative differences."""
    eel = eel or inputs.eel3d()
    out: Dict[str, dict] = {}
    us = {}
    for b in ("capsules", "mask"):
        u, drag, h = run(eel, "f32", steps, log=log, boundary=b, history=True)
        us[b] = u
        half = h[steps // 2:].sum(axis=1)
        out[b] = {"mean_drag": float(half[:, 0].mean()),
                  "lateral_amplitude": float(0.5 * (half[:, 1].max() - half[:, 1].min())),
                  "yaw_torque_amplitude": float(0.5 * (half[:, 5].max() - half[:, 5].min()))}
    ref = out["capsules"]
    rel = {k: abs(out["mask"][k] - ref[k]) / abs(ref[k]) for k in ref}
    u_rel = (torch.linalg.vector_norm(us["mask"] - us["capsules"]) / torch.linalg.vector_norm(us["capsules"])).item()
    del us
    torch.cuda.empty_cache()
    return {"workload": f"d3q27_eel_{eel.nx}x{eel.ny}x{eel.nz}", "steps": steps, "capsules": ref,
            "mask": out["mask"], "mask_rel_diff": rel, "u_rel_l2_mask_vs_capsules": u_rel}


def sweep(steps: int = 5000, dtypes: Iterable[str] = ("f64", "f32", "f16", "bf16"), eel=None, log=None) -> Dict:
    """Error table of each storage precision against the FP64 run."""
    eel = eel or inputs.eel3d()
    out: Dict[str, dict] = {}
    ref_u: Optional[torch.Tensor] = None
    ref_drag = None
    for dt in dtypes:
        u, drag = run(eel, dt, steps, log=log)
        half = drag[steps // 2:].mean().item()
        if dt == "f64":
            ref_u, ref_drag = u, half
            out[dt] = {"u_rel_l2": 0.0, "mean_drag": half, "mean_drag_rel_err": 0.0}
        else:
            if ref_u is None:
                raise ValueError("the FP64 run must come first")
            err = (torch.linalg.vector_norm(u - ref_u) / torch.linalg.vector_norm(ref_u)).item()
            out[dt] = {"u_rel_l2": err, "mean_drag": half, "mean_drag_rel_err": abs(half - ref_drag) / abs(ref_drag)}
        if log:
            log(f"[precision] {dt}: {out[dt]}")
        del u
        torch.cuda.empty_cache()
    return {"workload": f"d3q27_eel_{eel.nx}x{eel.ny}x{eel.nz}", "steps": steps, "re": 500.0,
            "segments": eel.nseg, "length": eel.length, "radius": eel.radius, "errors": out}

"""Precision sweep: the same D3Q27 run with moments stored in FP64 (the GPU
oracle: the gather kernel in FP64), FP32, FP16 and BF16 (16-bit storage of
rho - 1 and per-field scaled u and S, collision in FP32).

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


def run(eel: "inputs.Eel3D", dtype: str, steps: int, seed: int = 4004, re: float = 500.0, log=None):
    """Advance the eel workload `steps` steps in `dtype`; returns (final u as
    an FP64 device tensor [3, nz, ny, nx], drag history [steps] FP64)."""
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
        if dtype in ("f16", "bf16"):
            cur = f[1:4].float() / L.scale[1]
            f[1:4].copy_(((cur + noise) * L.scale[1]).to(f.dtype))
        else:
            f[1:4] += noise.to(f.dtype)
    caps, poses = eel.tables(0, steps)
    box = eel.box(0, steps)
    caps_d = torch.from_numpy(caps).to(L.device)
    poses_d = torch.from_numpy(poses).to(L.device)
    hist = torch.zeros((steps, nseg, 6), dtype=torch.float64, device=L.device)
    for t in range(steps):
        L.voxelize_device(caps_d[t], box)
        L.solids[:nseg].copy_(poses_d[t])
        L.step(1, force_history=hist[t:t + 1])
        if log and (t + 1) % 1000 == 0:
            log(f"[precision] {dtype}: step {t + 1}/{steps}")
    u = L.fields()[1:4].to(torch.float64)
    if dtype in ("f16", "bf16"):
        u = u / L.scale[1]
    drag = hist[:, :, 0].sum(1)
    if not bool(torch.isfinite(u).all()):
        raise RuntimeError(f"{dtype} run diverged")
    return u, drag


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

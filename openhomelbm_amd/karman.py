"""Karman vortex street diagnostics: drag/lift history of the D2Q9 cylinder
workload, Strouhal number from the lift spectrum, and the comparison of the
FP32 run against the FP64 GPU oracle run of the same inputs.

No reference counterpart: the reference records per-step forces inside its
environments (envs/lbm/karman/karman_lbm_env.py) but computes neither the
Strouhal number nor mean coefficients.
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np


def strouhal(lift: np.ndarray, diameter: float, velocity: float) -> float:
    """St = f D / U with f the dominant frequency (cycles per step) of the
    mean-removed lift series: the peak of the Hann-windowed spectrum, refined
    by a parabola through the log-magnitudes of the peak bin and its neighbours."""
    x = np.asarray(lift, dtype=np.float64)
    x = x - x.mean()
    n = len(x)
    if n < 16:
        raise ValueError("lift series too short")
    spec = np.abs(np.fft.rfft(x * np.hanning(n)))
    k = int(np.argmax(spec[1:])) + 1
    if 1 <= k < len(spec) - 1:
        a, b, c = np.log(spec[k - 1] + 1e-300), np.log(spec[k] + 1e-300), np.log(spec[k + 1] + 1e-300)
        den = a - 2 * b + c
        delta = 0.5 * (a - c) / den if den != 0 else 0.0
    else:
        delta = 0.0
    f = (k + delta) / n
    return f * diameter / velocity


def run(steps: int, dtype: str = "f32", top_bottom: int = 2, seed: int = 1001, log=None,
        boundary: str = "mask", segments: int = 96):
    """The 512x128 cylinder workload (disc D = 16 at (96, 65), u = 0.05,
    nu = 0.008) for `steps` steps; returns FP64 (drag, lift) histories.
    boundary "mask": the integer disc mask and the mask link rule;
    "cut_cell": the disc as a regular polygon of `segments` sides inscribed
    in the circle, with the reference's segment-link cut cells."""
    import torch

    from . import inputs

    w = inputs.d2q9_karman(seed)
    w.bc_type = (0, top_bottom, 1, top_bottom)
    lat = inputs.build_lattice(w, dtype)
    hist = torch.zeros((steps, 1, 6), dtype=torch.float64, device=lat.device)
    if boundary == "cut_cell":
        ang = 2 * np.pi * np.arange(segments) / segments
        poly = np.stack([96 + 8 * np.cos(ang), 65 + 8 * np.sin(ang)], 1)[None].astype(np.float32)
        verts = torch.from_numpy(poly).to(lat.device)
        counts = torch.tensor([segments], dtype=torch.int32, device=lat.device)
        radius2 = torch.tensor([(8 + 2 ** 0.5) ** 2], dtype=torch.float32, device=lat.device)
    elif boundary != "mask":
        raise ValueError("boundary must be 'mask' or 'cut_cell'")
    chunk = 5000
    for k0 in range(0, steps, chunk):
        k1 = min(steps, k0 + chunk)
        if boundary == "cut_cell":
            lat.step_polygons(k1 - k0, verts, verts, counts, radius2, force_history=hist[k0:k1])
        else:
            lat.step(k1 - k0, force_history=hist[k0:k1])
        if log:
            log(f"[karman] {dtype}: step {k1}/{steps}")
    h = hist[:, 0].cpu().numpy()
    if not np.isfinite(h).all():
        raise RuntimeError(f"{dtype} Karman run diverged")
    return h[:, 0], h[:, 1]


def compare(steps: int = 40000, window: int = 25000, log=None) -> Dict:
    """FP32 against the FP64 GPU oracle: Strouhal number, mean drag and lift
    amplitude over the last `window` steps, and their relative differences."""
    D, U = 16.0, 0.05
    out: Dict[str, dict] = {}
    for dt in ("f64", "f32"):
        drag, lift = run(steps, dt, log=log)
        d, l = drag[-window:], lift[-window:]
        out[dt] = {"strouhal": strouhal(l, D, U), "mean_drag": float(d.mean()),
                   "lift_amplitude": float(0.5 * (l.max() - l.min()))}
    ref = out["f64"]
    rel = {k: abs(out["f32"][k] - ref[k]) / abs(ref[k]) for k in ref}
    return {"workload": "d2q9_karman_512x128", "steps": steps, "window": window, "faces_top_bottom": "free slip",
            "re": 100, "f64": ref, "f32": out["f32"], "f32_rel_err": rel}


def compare_boundaries(steps: int = 40000, window: int = 25000, log=None) -> Dict:
    """FP32 runs of the cylinder with the mask link rule and with the
    reference's exact cut cells on a 96-segment polygon: Strouhal number,
    mean drag and lift amplitude, and the mask rule's relative difference."""
    D, U = 16.0, 0.05
    out: Dict[str, dict] = {}
    for b in ("cut_cell", "mask"):
        drag, lift = run(steps, "f32", log=log, boundary=b)
        d, l = drag[-window:], lift[-window:]
        out[b] = {"strouhal": strouhal(l, D, U), "mean_drag": float(d.mean()),
                  "lift_amplitude": float(0.5 * (l.max() - l.min()))}
    ref = out["cut_cell"]
    rel = {k: abs(out["mask"][k] - ref[k]) / abs(ref[k]) for k in ref}
    return {"workload": "d2q9_karman_512x128", "steps": steps, "window": window, "segments": 96,
            "cut_cell": ref, "mask": out["mask"], "mask_rel_diff": rel}

"""Autotuners: the tiled D3Q27 kernel's launch shape, and the batched
D2Q9 step's strip heights (tune_batch).

The tiled kernel's tile (TX x TY columns staged in LDS) and z-chunk (ZC
planes marched per workgroup) trade halo re-reads, LDS footprint (hence
workgroups per CU) and the number of workgroups in flight; the best choice
depends on the lattice size. The tuner times every supported shape on a
lattice of the size class being tuned (the state is saved and restored) and
records the fastest per size class in a JSON table. Lattice.step applies the
table's entry for its size class before it launches.

Results do not depend on the tile shape (the kernel's arithmetic per cell is
the same), so tuning changes speed only.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Iterable, Optional, Tuple

from . import _native as N

#: tile shapes the tiled kernel is compiled for
TILES = ((64, 8), (64, 4), (32, 8), (32, 16), (16, 16))
ZCHUNKS = (8, 16, 32, 64)
#: shipped table (measured on MI355X); a user table path may be set with
#: OPENHOMELBM_TUNING
DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")

#: the shape a lattice of an untuned size class runs with
DEFAULT_TILE = (64, 8, 16)

_table: Optional[Dict[str, Tuple[int, int, int]]] = None


def size_class(lattice: int, nx: int, ny: int, nz: int = 1, layout: str = "rows") -> str:
    """Size class key: lattice, cell count bucket (powers of 4 cells) and,
    for the structure-of-arrays D3Q27 layout, a ':soa' suffix."""
    n = nx * ny * nz
    b = 0
    while (1 << (2 * (b + 1))) <= n:
        b += 1
    suffix = ":soa" if lattice == 27 and layout == "soa" else ""
    return f"d{2 if lattice == 9 else 3}q{lattice}:4^{b}{suffix}"


def table_path() -> str:
    return os.environ.get("OPENHOMELBM_TUNING", DEFAULT_TABLE)


def load_table(path: Optional[str] = None) -> Dict[str, Tuple[int, int, int]]:
    global _table
    p = path or table_path()
    t: Dict[str, Tuple[int, int, int]] = {}
    if os.path.exists(p):
        with open(p) as f:
            raw = json.load(f)
        for k, v in raw.get("tiles", {}).items():
            t[k] = tuple(int(x) for x in v["tile"])
    _table = t
    return t


def save_table(entries: Dict[str, dict], path: Optional[str] = None) -> str:
    p = path or table_path()
    raw = {"tiles": {}}
    if os.path.exists(p):
        with open(p) as f:
            raw = json.load(f)
    raw.setdefault("tiles", {}).update(entries)
    with open(p, "w") as f:
        json.dump(raw, f, indent=1, sort_keys=True)
    load_table(p)
    return p


def tuned_tile(lattice: int, nx: int, ny: int, nz: int, layout: str = "rows") -> Optional[Tuple[int, int, int]]:
    if _table is None:
        load_table()
    return _table.get(size_class(lattice, nx, ny, nz, layout))


def apply(lat) -> None:
    """Set the tiled kernel's shape for this lattice's size class: the tuned
    entry, or DEFAULT_TILE for a class with none (so the shape never depends
    on which lattice stepped before)."""
    if lat.lattice != 27:
        return
    t = tuned_tile(lat.lattice, lat.nx, lat.ny, lat.nz, lat.layout) or DEFAULT_TILE
    if N.get_tiling(27) != tuple(t):
        N.set_tiling(27, *t)


def candidates() -> Iterable[Tuple[int, int, int]]:
    for tx, ty in TILES:
        for zc in ZCHUNKS:
            yield (tx, ty, zc)


def tune(lat, steps: int = 20, warmup: int = 3, shapes: Optional[Iterable[Tuple[int, int, int]]] = None,
         save: bool = True, path: Optional[str] = None, log=None, rounds: int = 2, settle: int = 200) -> dict:
    """Time the tiled kernel on `lat` for every shape; keep the fastest.

    `settle` steps run first (the clocks ramp up over the first launches);
    every shape is then timed `rounds` times, interleaved, and keeps its best
    time. The lattice's state is restored afterwards. Returns the size-class
    entry {"tile": [tx, ty, zc], "us_per_step": t, "timings": {...}}."""
    import torch

    if lat.lattice != 27:
        raise ValueError("only the D3Q27 kernel has tunable tile shapes")
    saved = lat.buf.clone()
    cur = lat.cur
    timings = {}
    shapes = list(shapes or candidates())
    try:
        if settle > 0:
            N.set_tiling(27, *shapes[0])
            lat.step(settle, N.HL_KERNEL_TILED, _tuned=False)
        for _ in range(max(rounds, 1)):
            for t in shapes:
                N.set_tiling(27, *t)
                lat.step(warmup, N.HL_KERNEL_TILED, _tuned=False)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                torch.cuda.synchronize()
                e0.record()
                lat.step(steps, N.HL_KERNEL_TILED, _tuned=False)
                e1.record()
                torch.cuda.synchronize()
                key = "x".join(map(str, t))
                us = e0.elapsed_time(e1) * 1e3 / steps
                timings[key] = min(us, timings.get(key, us))
                if log:
                    log(f"[autotune] tile {t}: {us:.1f} us/step")
    finally:
        lat.buf.copy_(saved)
        lat.cur = cur
    best = min(timings, key=timings.get)
    tile = [int(x) for x in best.split("x")]
    N.set_tiling(27, *tile)
    entry = {"tile": tile, "us_per_step": round(timings[best], 2),
             "timings": {k: round(v, 2) for k, v in timings.items()}}
    if save:
        save_table({size_class(lat.lattice, lat.nx, lat.ny, lat.nz, lat.layout): entry}, path)
    return entry


#: strip heights (rows, body_rows) the batch tuner tries
BATCH_STRIPS = ((16, 4), (12, 4), (24, 4), (16, 3), (16, 6), (8, 4), (32, 4))


def tune_batch(batch, steps: int = 20, reps: int = 3, shapes: Optional[Iterable[Tuple[int, int]]] = None) -> dict:
    """Time the batched step's strip heights (Batch.set_strips) on this batch
    by replays of a captured graph of `steps` steps, keep the fastest and
    return {"strips": best, "ms_per_step": {shape: ms}}. The batch's states,
    parity and time are restored afterwards; states do not depend on the
    heights (hl_batch_set_strips), so tuning changes speed only."""
    import torch

    lats = batch.lattices
    saved = [(L.fields(0).clone(), L.fields(1).clone()) for L in lats]
    par, tim = batch.pull_state()
    timings = {}
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for shape in (shapes or BATCH_STRIPS):
        batch.set_strips(*shape)
        batch.capture(steps)
        batch.replay()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            e0.record()
            batch.replay()
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1) / steps)
        timings[tuple(shape)] = round(best, 5)
    win = min(timings, key=timings.get)
    batch.set_strips(*win)
    for L, (a, b) in zip(lats, saved):
        L.fields(0).copy_(a)
        L.fields(1).copy_(b)
    for L, p in zip(lats, par):
        L.cur = p
    batch.time = list(tim)
    batch._push_state()
    torch.cuda.synchronize()
    return {"strips": win, "ms_per_step": {f"{r}x{h}": v for (r, h), v in timings.items()}}


def main(argv=None) -> None:
    """Tune the size classes of the shipped workloads and write the table:
    python -m openhomelbm_amd.autotune [--out PATH] [--steps K]."""
    import argparse
    import sys

    from .lattice import Lattice

    p = argparse.ArgumentParser()
    p.add_argument("--out", default=None, help="table path (default: the shipped table)")
    p.add_argument("--steps", type=int, default=50)
    a = p.parse_args(argv)
    # one lattice per size class: 96^3 (4^9), 128^3 (4^10), the 256x128x128
    # sphere (4^11)
    for shape in ((96, 96, 96), (128, 128, 128), (256, 128, 128)):
        lat = Lattice(27, *shape)
        lat.nu = 0.01
        lat.bc_type = [0, 1, 1, 1, 1, 1]
        lat.bc_value = [[0.05, 0.0, 0.0]] + [[0.0, 0.0, 0.0]] * 5
        lat.init_uniform((0.05, 0.0, 0.0))
        e = tune(lat, steps=a.steps, path=a.out, log=lambda m: print(m, file=sys.stderr, flush=True))
        print(f"{size_class(27, *shape)}: {e['tile']} {e['us_per_step']} us", flush=True)
        del lat


if __name__ == "__main__":
    main()

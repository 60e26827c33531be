"""Serving loop: 'advance environment e by k steps and return its body forces'
requests from many resident swimmer environments, with continuous batching.

* Every admitted environment keeps its lattice resident on the GPU. An
  environment is admitted when its HBM footprint fits the budget; otherwise it
  waits until released environments free enough memory.
* The loop advances in ticks. Each tick takes every environment whose
  oldest request still has steps left and advances all of them by m steps:
  the fewest steps any of those requests still needs (at most
  max_tick_steps, and within each environment's window of kinematics
  tables):
  * all such D2Q9 environments advance m steps in one call of the batched
    engine (Batch.step with an active mask and an m-step device force
    history);
  * each such D3Q27 environment advances m steps on one of two side
    streams, so their kernels overlap on the device.
  A request that arrives while others are running joins at the next tick.
* A request completes when its k-th step has run. Its result is the FP64
  force/torque history of its k steps, [k, nsolid, 6], gathered on the
  device (one copy per environment and tick) and read back once. Latency
  runs from arrival to when the host sees the completion event.
* Results do not depend on what else is in the batch. The batched D2Q9
  kernels give every environment the same arithmetic as stepping it alone, and
  each D3Q27 environment runs its own kernels, so a request's result is
  bitwise the same alone and in the mix.

There is no reference counterpart: the reference steps every world on every
call (lbm_solver.py:97-141).
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import Deque, Dict, List, Optional, Sequence

import numpy as np
import torch

from . import inputs
from .batch import MAX_SOLIDS, Batch
from .lattice import Lattice


@dataclass
class EnvSpec:
    """One swimmer environment: a D2Q9 eel (inputs.Eel) or a D3Q27 eel (inputs.Eel3D)."""

    eel: object
    lattice: int
    horizon: int = 4096  # steps of kinematics tables kept on the device
    seed: int = 0

    @property
    def shape(self):
        e = self.eel
        return (e.nx, e.ny) if self.lattice == 9 else (e.nx, e.ny, e.nz)

    @property
    def cells(self) -> int:
        return int(np.prod(self.shape))

    @property
    def nsolid(self) -> int:
        return inputs.EEL_SEGMENTS if self.lattice == 9 else self.eel.nseg

    def hbm_bytes(self) -> int:
        """Device footprint: two moment states, mask, kinematics tables, forces."""
        nf = 6 if self.lattice == 9 else 10
        cells = (self.cells + 63) // 64 * 64
        pose = 8 if self.lattice == 9 else 24
        return cells * (nf * 4 * 2 + 1) + self.horizon * self.nsolid * (8 + pose) * 4 + self.horizon * MAX_SOLIDS * 48


@dataclass
class Request:
    env: int
    steps: int
    arrival: float
    remaining: int = 0
    done: bool = False
    finish: float = 0.0
    forces: Optional[np.ndarray] = None
    _dev: Optional[torch.Tensor] = field(default=None, repr=False)  # [steps, nsolid, 6] on the device
    _off: int = 0

    @property
    def latency(self) -> float:
        return self.finish - self.arrival


class _Resident:
    def __init__(self, spec: EnvSpec, device, max_tick: int = 16):
        self.spec = spec
        e = spec.eel
        n = spec.nsolid
        if spec.lattice == 9:
            L = Lattice(9, e.nx, e.ny, nsolid=n, device=device)
            L.nu = e.nu
            L.bc_type = [0, 2, 1, 2, 1, 1]
            L.set_state(inputs.eel_state(e, spec.seed))
        else:
            L = Lattice(27, e.nx, e.ny, e.nz, nsolid=n, device=device)
            L.nu = e.nu()
            L.bc_type = [0, 1, 2, 2, 2, 2]
            L.init_uniform((inputs.EEL_U, 0.0, 0.0))
        L.bc_value = [[inputs.EEL_U, 0.0, 0.0]] + [[0.0, 0.0, 0.0]] * 5
        self.lat = L
        self.time = 0
        self.refresh(0)
        if spec.lattice == 27:
            L.mask = torch.zeros(L.ncell, dtype=torch.uint8, device=L.device)
            L.mask_changed()
            self.hist = torch.zeros((max_tick, 1, n, 6), dtype=torch.float64, device=L.device)
        self.queue: Deque[Request] = collections.deque()

    def refresh(self, t0: int) -> None:
        """Kinematics tables for steps t0 .. t0+horizon-1, stored so that step t
        is at index t % horizon (the batch reads its tables that way)."""
        H = self.spec.horizon
        caps, poses = self.spec.eel.tables(t0, H)
        self.caps_np = np.roll(caps, t0 % H, axis=0)
        self.poses_np = np.roll(poses, t0 % H, axis=0)
        self.box = self.spec.eel.box(t0, H)
        self.caps = torch.from_numpy(self.caps_np).to(self.lat.device)
        self.poses = torch.from_numpy(self.poses_np).to(self.lat.device)
        self.t_end = t0 + H


class Server:
    """Continuous-batching request loop over resident environments.

    Args:
        hbm_budget: bytes of device memory the resident environments may use.
        streams: side streams over which D3Q27 environments are spread per
            tick (default 2: with the caller's stream that stays within the
            four hardware queues HIP gives a process).
        max_tick_steps: the most steps one tick advances.
    """

    def __init__(self, hbm_budget: float = 200e9, streams: int = 2, device=None, max_tick_steps: int = 16):
        self.max_tick = int(max_tick_steps)
        if self.max_tick < 1:
            raise ValueError("max_tick_steps must be >= 1")
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.budget = float(hbm_budget)
        self.used = 0
        self.envs: Dict[int, _Resident] = {}
        self.pending: Deque[tuple] = collections.deque()  # (id, spec) waiting for memory
        self._next_id = 0
        self._batch: Optional[Batch] = None
        self._batch_ids: List[int] = []
        self._streams = [torch.cuda.Stream(device=self.device) for _ in range(max(streams, 1))]
        self._inflight: Deque[tuple] = collections.deque()  # (event, [requests completed by it])
        self.cells_stepped = 0
        self.ticks = 0
        self.requests: List[Request] = []

    # -------------------------------------------------------- admission
    def register(self, spec: EnvSpec) -> int:
        """Ask for an environment; it becomes resident when it fits the budget."""
        if spec.hbm_bytes() > self.budget:
            raise ValueError(f"environment needs {spec.hbm_bytes()} bytes, more than the whole budget")
        eid = self._next_id
        self._next_id += 1
        self.pending.append((eid, spec))
        self._admit()
        return eid

    def _admit(self):
        admitted = False
        while self.pending and self.used + self.pending[0][1].hbm_bytes() <= self.budget:
            eid, spec = self.pending.popleft()
            self.envs[eid] = _Resident(spec, self.device, self.max_tick)
            self.used += spec.hbm_bytes()
            admitted = True
        if admitted:
            self._rebuild_batch()

    def release(self, eid: int) -> None:
        """Free an environment (its queued requests must be complete)."""
        r = self.envs.get(eid)
        if r is None:
            raise KeyError(eid)
        if r.queue:
            raise RuntimeError("environment still has requests")
        self.drain()
        del self.envs[eid]
        self.used -= r.spec.hbm_bytes()
        self._rebuild_batch()
        self._admit()

    def resident(self, eid: int) -> bool:
        return eid in self.envs

    def _rebuild_batch(self):
        ids = [i for i, r in self.envs.items() if r.spec.lattice == 9]
        if ids == self._batch_ids:
            return
        self.drain()
        if self._batch is not None:
            self._batch.close()
        self._batch, self._batch_ids = None, ids
        if not ids:
            return
        b = Batch([self.envs[i].lat for i in ids])
        for k, i in enumerate(ids):
            r = self.envs[i]
            b.set_kinematics(k, r.caps_np, r.poses_np, r.box)
            b.time[k] = r.time
        b._push_state()
        self._batch = b
        self._bhist = b.history_buffer(self.max_tick)

    # --------------------------------------------------------- requests
    def submit(self, eid: int, steps: int, arrival: Optional[float] = None) -> Request:
        if steps < 1:
            raise ValueError("steps must be >= 1")
        if eid not in self.envs and all(p[0] != eid for p in self.pending):
            raise KeyError(eid)
        spec = self.envs[eid].spec if eid in self.envs else next(p[1] for p in self.pending if p[0] == eid)
        req = Request(env=eid, steps=steps, arrival=time.perf_counter() if arrival is None else arrival,
                      remaining=steps)
        req._dev = torch.empty((steps, spec.nsolid, 6), dtype=torch.float64, device=self.device)
        self.requests.append(req)
        if eid in self.envs:
            self.envs[eid].queue.append(req)
        else:
            self._waiting = getattr(self, "_waiting", collections.defaultdict(collections.deque))
            self._waiting[eid].append(req)
        return req

    def _ready(self) -> List[int]:
        waiting = getattr(self, "_waiting", None)
        if waiting:
            for eid in list(waiting):
                if eid in self.envs:
                    self.envs[eid].queue.extend(waiting.pop(eid))
        return [i for i, r in self.envs.items() if r.queue]

    # ------------------------------------------------------------- ticks
    def tick(self) -> int:
        """Advance every environment that has work by the tick's m steps;
        returns how many environments advanced."""
        ready = self._ready()
        if not ready:
            return 0
        cur = torch.cuda.current_stream(self.device)
        for i in ready:
            r = self.envs[i]
            if r.time >= r.t_end:  # next window of the kinematics tables
                r.refresh(r.time)
                if r.spec.lattice == 9:
                    self._batch.set_kinematics(self._batch_ids.index(i), r.caps_np, r.poses_np, r.box)
        m = min([self.max_tick] + [self.envs[i].queue[0].remaining for i in ready] +
                [self.envs[i].t_end - self.envs[i].time for i in ready])
        done_now: List[Request] = []
        # D2Q9: m steps of every ready environment in one batched call
        ready2 = [i for i in ready if self.envs[i].spec.lattice == 9]
        if ready2:
            active = np.zeros(len(self._batch_ids), bool)
            for i in ready2:
                active[self._batch_ids.index(i)] = True
            self._batch.step(m, active=active, force_history=self._bhist, stream=cur)
            for i in ready2:
                r = self.envs[i]
                req = r.queue[0]
                k = self._batch_ids.index(i)
                req._dev[req._off:req._off + m].copy_(self._bhist[:m, k, : r.spec.nsolid])
                self.cells_stepped += r.spec.cells * m
        # D3Q27: per environment, m steps, spread over the side streams
        ready3 = [i for i in ready if self.envs[i].spec.lattice == 27]
        if ready3:
            fork = torch.cuda.Event()
            fork.record(cur)
            for n, i in enumerate(ready3):
                r = self.envs[i]
                req = r.queue[0]
                s = self._streams[n % len(self._streams)]
                s.wait_event(fork)
                with torch.cuda.stream(s):
                    for q in range(m):
                        j = (r.time + q) % r.spec.horizon
                        r.lat.voxelize_device(r.caps[j], r.box, stream=s)
                        r.lat.solids[: r.spec.nsolid].copy_(r.poses[j])
                        r.lat.step(1, force_history=r.hist[q], stream=s)
                    req._dev[req._off:req._off + m].copy_(r.hist[:m, 0])
                req._dev.record_stream(s)
                self.cells_stepped += r.spec.cells * m
            for s in self._streams[: len(ready3)]:
                j = torch.cuda.Event()
                j.record(s)
                cur.wait_event(j)
        for i in ready:
            r = self.envs[i]
            r.time += m
            req = r.queue[0]
            req._off += m
            req.remaining -= m
            if req.remaining == 0:
                r.queue.popleft()
                done_now.append(req)
        ev = torch.cuda.Event()
        ev.record(cur)
        self._inflight.append((ev, done_now))
        self.ticks += 1
        self._poll(block=len(self._inflight) > 2)
        return len(ready)

    def _finish(self, reqs):
        t = time.perf_counter()
        for req in reqs:
            req.forces = req._dev.cpu().numpy()
            req._dev = None
            req.finish = t
            req.done = True

    def _poll(self, block: bool = False):
        while self._inflight:
            ev, reqs = self._inflight[0]
            if block:
                ev.synchronize()
                block = len(self._inflight) > 3
            elif not ev.query():
                return
            self._inflight.popleft()
            self._finish(reqs)

    def drain(self):
        """Wait for every enqueued tick and complete its requests."""
        while self._inflight:
            ev, reqs = self._inflight.popleft()
            ev.synchronize()
            self._finish(reqs)

    def run_until_idle(self):
        while self.tick():
            pass
        self.drain()

    # ----------------------------------------------------------- driver
    def serve(self, arrivals: Sequence[tuple], log=None) -> dict:
        """Replay (time offset s, env id, steps) arrivals in real time; returns stats."""
        arrivals = sorted(arrivals)
        t0 = time.perf_counter()
        k = 0
        reqs = []
        while k < len(arrivals) or any(r.queue for r in self.envs.values()) or self._inflight or \
                any(getattr(self, "_waiting", {}).values()):
            now = time.perf_counter() - t0
            while k < len(arrivals) and arrivals[k][0] <= now:
                off, eid, steps = arrivals[k]
                reqs.append(self.submit(eid, steps, arrival=t0 + off))
                k += 1
            if not self.tick():
                self._poll(block=bool(self._inflight))
                if k < len(arrivals) and not self._inflight:
                    time.sleep(max(0.0, min(arrivals[k][0] - (time.perf_counter() - t0), 1e-3)))
        self.drain()
        wall = time.perf_counter() - t0
        lat = np.array([r.latency for r in reqs]) * 1e3
        return {"requests": len(reqs), "wall_s": round(wall, 3), "ticks": self.ticks,
                "mlups": round(self.cells_stepped / wall / 1e6, 1),
                "latency_ms_p50": round(float(np.percentile(lat, 50)), 3) if len(lat) else None,
                "latency_ms_p99": round(float(np.percentile(lat, 99)), 3) if len(lat) else None}


def pool(n: int = 32, seed: int = 5005, horizon: int = 4096, frac2d: float = 0.7) -> List[EnvSpec]:
    """Mixed environment pool: 70% D2Q9 (128x64 ... 512x256), 30% D3Q27
    (96^3 ... 192^3), each a seeded CPG capsule chain of length 0.3 nx."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes2 = ((128, 64), (256, 128), (384, 192), (512, 256))
    sizes3 = (96, 128, 160, 192)
    # exactly round(frac2d * n) D2Q9 environments, in a seeded order
    kinds = np.array([True] * int(round(frac2d * n)) + [False] * (n - int(round(frac2d * n))))
    rng.shuffle(kinds)
    out = []
    for k in range(n):
        A, f, phi = float(rng.uniform(0.2, 0.6)), float(rng.uniform(0.5, 2.0)), float(rng.uniform(0, np.pi / 2))
        if kinds[k]:
            nx, ny = sizes2[int(rng.integers(0, 4))]
            e = inputs.Eel(nx=nx, ny=ny, amplitude=A, frequency=f, phase_lag=phi, head=(0.25 * nx, 0.5 * ny),
                           seg_len=0.3 * nx / inputs.EEL_SEGMENTS)
            lat = 9
        else:
            s = sizes3[int(rng.integers(0, 4))]
            e = inputs.Eel3D(nx=s, ny=s, nz=s, amplitude=A, frequency=f, phase_lag=phi, length=0.3 * s,
                             radius=max(2.0, s / 48.0), nseg=8)
            lat = 27
        v = e.max_speed()
        if v > inputs.EEL_VMAX:
            e.time_scale = inputs.EEL_VMAX / v
        out.append(EnvSpec(eel=e, lattice=lat, horizon=horizon, seed=k))
    return out


def poisson_arrivals(rate: float, duration: float, env_ids: Sequence[int], seed: int = 5005,
                     ks: Sequence[int] = (4, 8, 16)) -> List[tuple]:
    """(time, env, k) with exponential gaps of mean 1/rate over `duration` seconds."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t, out = 0.0, []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > duration:
            return out
        out.append((t, int(rng.choice(env_ids)), int(rng.choice(ks))))


def sweep(rates: Sequence[float] = (250.0, 500.0, 1000.0, 2000.0, 4000.0), duration: float = 1.0, n_envs: int = 32,
          hbm_budget: float = 200e9, seed: int = 5005, log=None) -> dict:
    """Serve Poisson request streams at each rate over one resident pool;
    report throughput and p50/p99 request latency per rate."""
    specs = pool(n_envs, seed=seed, horizon=4096)
    srv = Server(hbm_budget=hbm_budget)
    ids = [srv.register(s) for s in specs]
    warm = [srv.submit(i, 2) for i in ids]  # first launches of every kernel
    srv.run_until_idle()
    out = []
    for k, rate in enumerate(rates):
        srv.cells_stepped = 0
        srv.ticks = 0
        srv.requests = []
        st = srv.serve(poisson_arrivals(rate, duration, ids, seed=seed + k))
        st["rate_per_s"] = rate
        out.append(st)
        if log:
            log(f"[serving] {st}")
    n2 = sum(1 for s in specs if s.lattice == 9)
    return {"envs": n_envs, "d2q9": n2, "d3q27": n_envs - n2, "hbm_budget_gb": hbm_budget / 1e9,
            "resident_gb": round(srv.used / 1e9, 3), "k": [4, 8, 16], "duration_s": duration, "rates": out,
            "warmup_requests": len(warm)}

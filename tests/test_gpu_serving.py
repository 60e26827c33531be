"""The serving loop: admission under an HBM budget, continuous batching of
D2Q9 and D3Q27 environments, and results that are the same alone and in the
mix (bitwise: every environment's arithmetic is independent of the batch)."""
import numpy as np
import pytest
import torch

from openhomelbm_amd import inputs, serving

pytestmark = pytest.mark.gpu


def _specs():
    e2a = inputs.Eel(nx=96, ny=48, amplitude=0.4, frequency=1.5, phase_lag=0.6, head=(24, 24), seg_len=3.6)
    e2b = inputs.Eel(nx=130, ny=40, amplitude=0.3, frequency=1.0, phase_lag=0.9, head=(32, 20), seg_len=4.8)
    e3 = inputs.Eel3D(nx=40, ny=36, nz=34, amplitude=0.4, frequency=1.2, phase_lag=0.7, length=12.0, radius=2.5, nseg=4)
    for e in (e2a, e2b, e3):
        e.time_scale = min(1.0, inputs.EEL_VMAX / e.max_speed())
    return [serving.EnvSpec(e2a, 9, horizon=64, seed=1), serving.EnvSpec(e2b, 9, horizon=64, seed=2),
            serving.EnvSpec(e3, 27, horizon=64, seed=3)]


def test_request_alone_equals_in_mix():
    specs = _specs()
    # in the mix: all three environments, overlapping requests
    mix = serving.Server(hbm_budget=1e9)
    ids = [mix.register(s) for s in specs]
    r_other = mix.submit(ids[1], 16)
    mix.tick()
    mix.tick()
    r2 = mix.submit(ids[0], 8)     # joins while env 1 is running
    r3 = mix.submit(ids[2], 8)
    mix.run_until_idle()
    assert r_other.done and r2.done and r3.done
    # alone: a fresh server per environment
    for spec, r in ((specs[0], r2), (specs[2], r3)):
        solo = serving.Server(hbm_budget=1e9)
        i = solo.register(spec)
        q = solo.submit(i, 8)
        solo.run_until_idle()
        assert q.forces.shape == (8, spec.nsolid, 6)
        assert np.array_equal(q.forces, r.forces)
        assert torch.equal(solo.envs[i].lat.fields(), mix.envs[ids[specs.index(spec)]].lat.fields())
    assert np.abs(r2.forces[:, :, 0]).sum() > 0 and np.abs(r3.forces[:, :, 0]).sum() > 0


def test_admission_waits_for_memory():
    specs = _specs()
    budget = specs[1].hbm_bytes() + specs[2].hbm_bytes() + 1  # b is larger than a
    srv = serving.Server(hbm_budget=budget)
    a = srv.register(specs[0])
    c = srv.register(specs[2])
    b = srv.register(specs[1])          # does not fit yet
    assert srv.resident(a) and srv.resident(c) and not srv.resident(b)
    rb = srv.submit(b, 4)               # queued until admitted
    ra = srv.submit(a, 4)
    srv.run_until_idle()
    assert ra.done and not rb.done
    srv.release(a)
    assert srv.resident(b)
    srv.run_until_idle()
    assert rb.done and rb.forces.shape == (4, inputs.EEL_SEGMENTS, 6)
    with pytest.raises(ValueError):
        serving.Server(hbm_budget=10).register(specs[0])


def test_serve_poisson_reports_latency():
    specs = _specs()
    srv = serving.Server(hbm_budget=1e9)
    ids = [srv.register(s) for s in specs]
    arr = serving.poisson_arrivals(rate=400.0, duration=0.05, env_ids=ids, seed=3)
    assert len(arr) > 5
    st = srv.serve(arr)
    assert st["requests"] == len(arr)
    assert st["latency_ms_p50"] > 0 and st["latency_ms_p99"] >= st["latency_ms_p50"]
    assert all(r.done for r in srv.requests)


def test_concurrent_streams_do_not_share_scratch():
    """Two D3Q27 lattices of different sizes stepped on two streams at once
    (the near-body list and force partials are per caller stream) equal the
    same steps run one after the other."""
    from openhomelbm_amd.lattice import Lattice

    def make(nx, ny, nz, seed):
        L = Lattice(27, nx, ny, nz, nsolid=1)
        L.nu = 0.02
        L.bc_type = [0, 1, 1, 1, 1, 1]
        L.bc_value = [[0.05, 0, 0]] + [[0, 0, 0]] * 5
        p = np.zeros((1, 24), np.float32)
        p[0, 0:3] = p[0, 12:15] = (nx * 0.4, ny * 0.5, nz * 0.5)
        p[0, 3:12] = p[0, 15:24] = np.eye(3).ravel()
        L.set_poses(p)
        L.voxelize(np.array([[nx * 0.3, ny * 0.5, nz * 0.5, nx * 0.5, ny * 0.5, nz * 0.5, 4.0, 0]], np.float32))
        from tests.helpers import random_state
        L.set_state(random_state(27, (nz, ny, nx), seed=seed).astype(np.float32))
        return L

    shapes = ((72, 40, 36), (40, 34, 30))
    seq = [make(*s, seed=k) for k, s in enumerate(shapes)]
    hs = [torch.zeros((3, 1, 6), dtype=torch.float64, device="cuda") for _ in seq]
    for L, h in zip(seq, hs):
        L.step(3, force_history=h)
    torch.cuda.synchronize()
    par = [make(*s, seed=k) for k, s in enumerate(shapes)]
    hp = [torch.zeros((3, 1, 6), dtype=torch.float64, device="cuda") for _ in par]
    streams = [torch.cuda.Stream() for _ in par]
    for k in range(3):
        for L, h, s in zip(par, hp, streams):
            with torch.cuda.stream(s):
                L.step(1, force_history=h[k:k + 1], stream=s)
    torch.cuda.synchronize()
    for a, b, ha, hb in zip(seq, par, hs, hp):
        assert torch.equal(a.fields(), b.fields())
        assert torch.equal(ha, hb)

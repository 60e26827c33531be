"""The serving workload's pool at its own sizes: 70% D2Q9 environments
(128x64 ... 512x256) and 30% D3Q27 (96^3 ... 192^3), seeded capsule chains,
requests of k in {4, 8, 16} steps under an HBM budget that holds only part of
the pool, so admission waits for releases. Every request's forces and the
environment's state equal the same request served alone (bitwise), and one
D2Q9 and one D3Q27 environment of the mix follow the oracle over their first
request."""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from openhomelbm_amd import inputs, serving
from tests.helpers import tensor_rel_l2

pytestmark = pytest.mark.gpu

N_ENVS = 10
HORIZON = 64


def _pool():
    specs = serving.pool(N_ENVS, seed=5005, horizon=HORIZON)
    kinds = [s.lattice for s in specs]
    assert kinds.count(9) == 7 and kinds.count(27) == 3
    for s in specs:
        if s.lattice == 9:
            assert 128 <= s.eel.nx <= 512 and 64 <= s.eel.ny <= 256
        else:
            assert 96 <= s.eel.nx <= 192 and s.eel.nx == s.eel.ny == s.eel.nz
    return specs


def _ks():
    rng = np.random.Generator(np.random.PCG64(55))
    return [int(rng.choice((4, 8, 16))) for _ in range(N_ENVS)]


def test_pool_requests_alone_equal_in_mix_under_budget():
    specs, ks = _pool(), _ks()
    total = sum(s.hbm_bytes() for s in specs)
    budget = max(0.45 * total, max(s.hbm_bytes() for s in specs) * 1.01)
    srv = serving.Server(hbm_budget=budget)
    ids = [srv.register(s) for s in specs]
    waiting = [i for i in ids if not srv.resident(i)]
    assert waiting, "the budget must make some environments wait"
    assert srv.used <= budget
    reqs = [srv.submit(i, k) for i, k in zip(ids, ks)]
    done_envs = set()
    rounds = 0
    while not all(r.done for r in reqs):
        srv.run_until_idle()
        for i, r in zip(ids, reqs):
            if r.done and i not in done_envs and srv.resident(i):
                done_envs.add(i)
                states = srv.envs[i].lat.get_state(np.float32)
                r.final_state = states
                srv.release(i)  # frees memory: waiting environments are admitted
        assert srv.used <= budget
        rounds += 1
        assert rounds < 50
    for spec, r, k in zip(specs, reqs, ks):
        solo = serving.Server(hbm_budget=spec.hbm_bytes() * 1.01)
        i = solo.register(spec)
        q = solo.submit(i, k)
        solo.run_until_idle()
        assert q.forces.shape == (k, spec.nsolid, 6)
        assert np.array_equal(q.forces, r.forces)
        assert np.array_equal(solo.envs[i].lat.get_state(np.float32), r.final_state)
        del solo
        torch.cuda.empty_cache()


def _oracle_run(spec, nsteps):
    """The environment's first nsteps in the oracle: its initial state, per-step
    capsule masks and pose tables; returns the states and per-step forces."""
    e = spec.eel
    caps, poses = e.tables(0, nsteps)
    box = e.box(0, spec.horizon)
    if spec.lattice == 9:
        env = orc.Env(9, e.nx, e.ny, nu=e.nu, bc_type=(0, 2, 1, 2),
                      bc_value=((inputs.EEL_U, 0.0, 0.0),) + ((0.0, 0.0, 0.0),) * 5)
        m = inputs.eel_state(e, spec.seed)
    else:
        env = orc.Env(27, e.nx, e.ny, e.nz, nu=e.nu(), bc_type=(0, 1, 2, 2, 2, 2),
                      bc_value=((inputs.EEL_U, 0.0, 0.0),) + ((0.0, 0.0, 0.0),) * 5)
        m = orc.init_uniform(env, (inputs.EEL_U, 0.0, 0.0), "f32")
    forces = []
    for t in range(nsteps):
        mask = np.zeros(env.shape, np.uint8)
        orc.voxelize(env, caps[t], box, mask)
        env.mask = mask
        env.solids = poses[t]
        forces.append(orc.compute_forces(env, m))
        m = orc.step(env, m, "f32", omp=True)
    return m, np.array(forces)


@pytest.mark.parametrize("lattice", [9, 27])
def test_pool_environment_follows_oracle(lattice):
    """The first pool environment of each lattice, served in the mix for 4
    steps: D2Q9 bitwise (the batched kernels run the reference's operation
    order), D3Q27 within the tiled kernel's bar (relative L2 1e-6 per step);
    forces against the oracle's FP64 sums."""
    specs = _pool()
    k = [i for i, s in enumerate(specs) if s.lattice == lattice][0]
    srv = serving.Server(hbm_budget=200e9)
    ids = [srv.register(s) for s in specs]
    reqs = [srv.submit(i, 4) for i in ids]
    srv.run_until_idle()
    got = srv.envs[ids[k]].lat.get_state(np.float32)
    ref, fref = _oracle_run(specs[k], 4)
    F = reqs[k].forces
    if lattice == 9:
        assert np.array_equal(got, ref)
        np.testing.assert_allclose(F, fref, rtol=1e-10, atol=1e-12 * np.abs(fref).max())
    else:
        err = tensor_rel_l2(27, got, ref)["all"]
        assert err < 4e-6, err
        np.testing.assert_allclose(F[0], fref[0], rtol=1e-10, atol=1e-12 * np.abs(fref).max())
        np.testing.assert_allclose(F, fref, rtol=1e-4, atol=1e-6 * np.abs(fref).max())
    assert np.abs(fref[:, :, 0]).sum() > 0

"""GPU parity of the batched D2Q9 engine (hl_batch_*).

Environments of uneven sizes (none a multiple of the 62-column step strip) carry
8-capsule eels whose masks are re-voxelized on the device every step from
per-step capsule and pose tables. Checks:

* each environment's state after k steps equals the FP32 oracle stepped with
  the same per-step mask and poses, bit for bit, and the per-step force
  history agrees with the oracle's FP64 sums to summation-order rounding;
* the batched schedule equals the one-launch-per-environment schedule and
  stepping each lattice alone, bit for bit;
* inactive environments are untouched; a captured graph replayed several
  times equals eager stepping.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from openhomelbm_amd.batch import MAX_SOLIDS, Batch
from openhomelbm_amd.inputs import EEL_SEGMENTS, Eel, eel_state
from openhomelbm_amd.lattice import Lattice

pytestmark = pytest.mark.gpu

SIZES = ((72, 38), (100, 46), (130, 35))
BC_TYPE = (0, 2, 1, 2)
BC_VALUE = ((0.05, 0.0, 0.0),) + ((0.0, 0.0, 0.0),) * 5


def _eels(seed=7):
    rng = np.random.default_rng(seed)
    eels = []
    for nx, ny in SIZES:
        e = Eel(nx=nx, ny=ny, amplitude=float(rng.uniform(0.3, 0.6)), frequency=float(rng.uniform(1, 2)),
                phase_lag=float(rng.uniform(0.2, 1.2)), head=(0.25 * nx, 0.5 * ny), seg_len=0.3 * nx / EEL_SEGMENTS,
                radius=3.0)
        e.time_scale = min(1.0, 0.1 / e.max_speed())
        eels.append(e)
    return eels


def _lattice(e, index):
    L = Lattice(9, e.nx, e.ny, nsolid=EEL_SEGMENTS)
    L.nu = 0.02
    L.bc_type = list(BC_TYPE) + [1, 1]
    L.bc_value = [list(v) for v in BC_VALUE]
    L.set_state(eel_state(e, index, seed=11))
    return L


def _setup(t0=40, T=12):
    eels = _eels()
    lats = [_lattice(e, i) for i, e in enumerate(eels)]
    b = Batch(lats)
    tabs = []
    for i, e in enumerate(eels):
        caps, poses = e.tables(0, t0 + T)
        box = e.box(0, t0 + T)
        b.set_kinematics(i, caps, poses, box)
        b.set_time(i, t0)
        tabs.append((caps, poses, box))
    return eels, lats, b, tabs


def _oracle(e, m0, caps, poses, box, t0, nsteps):
    env = orc.Env(9, e.nx, e.ny, nu=0.02, bc_type=BC_TYPE, bc_value=BC_VALUE)
    m = m0
    hist = []
    for k in range(nsteps):
        t = t0 + k
        mask = np.zeros((e.ny, e.nx), np.uint8)
        orc.voxelize(env, caps[t], box, mask)
        env.mask = mask
        env.solids = poses[t]
        hist.append(orc.compute_forces(env, m))
        m = orc.step(env, m)
    return m, np.array(hist)


@pytest.mark.parametrize("nsteps", [1, 5])
def test_batch_matches_oracle(nsteps):
    t0 = 40
    eels, lats, b, tabs = _setup(t0)
    m0 = [L.get_state(np.float32) for L in lats]
    h = b.history_buffer(nsteps)
    b.step(nsteps, force_history=h)
    torch.cuda.synchronize()
    hist = h.cpu().numpy()
    par, tim = b.pull_state()
    assert par == [L.cur for L in lats] and tim == [t0 + nsteps] * len(lats)
    for i, (e, L) in enumerate(zip(eels, lats)):
        caps, poses, box = tabs[i]
        ref, ref_hist = _oracle(e, m0[i], caps, poses, box, t0, nsteps)
        got = L.get_state(np.float32)
        assert np.array_equal(got, ref), f"env {i} state differs from the FP32 oracle"
        F = hist[:, i, :EEL_SEGMENTS]
        scale = np.abs(ref_hist).max()
        assert scale > 0
        np.testing.assert_allclose(F, ref_hist, rtol=1e-10, atol=1e-12 * scale)
        assert not hist[:, i, EEL_SEGMENTS:].any()


def test_mask_follows_kinematics():
    t0 = 40
    eels, lats, b, tabs = _setup(t0)
    b.step(3)
    torch.cuda.synchronize()
    for i, (e, L) in enumerate(zip(eels, lats)):
        caps, _, box = tabs[i]
        env = orc.Env(9, e.nx, e.ny)
        want = np.zeros((e.ny, e.nx), np.uint8)
        orc.voxelize(env, caps[t0 + 2], box, want)
        got = L.mask.cpu().numpy().reshape(e.ny, e.nx)
        assert np.array_equal(got, want)
        assert set(np.unique(got)) == set(range(EEL_SEGMENTS + 1))


def test_batched_equals_per_env_and_alone():
    nsteps = 4
    _, lats_a, ba, _ = _setup()
    _, lats_b, bb, _ = _setup()
    ha, hb = ba.history_buffer(nsteps), bb.history_buffer(nsteps)
    ba.step(nsteps, force_history=ha)
    bb.step(nsteps, force_history=hb, per_env=True)
    torch.cuda.synchronize()
    assert torch.equal(ha, hb)
    for La, Lb in zip(lats_a, lats_b):
        assert torch.equal(La.fields(), Lb.fields())
    # one environment alone: a batch of one
    eels, lats_c, _, tabs = _setup()
    solo = Batch([lats_c[1]])
    caps, poses, box = tabs[1]
    solo.set_kinematics(0, caps, poses, box)
    solo.set_time(0, 40)
    hc = solo.history_buffer(nsteps)
    solo.step(nsteps, force_history=hc)
    torch.cuda.synchronize()
    assert torch.equal(lats_c[1].fields(), lats_a[1].fields())
    assert torch.equal(hc[:, 0], ha[:, 1])


def test_inactive_environment_untouched():
    _, lats, b, _ = _setup()
    before = [L.fields().clone() for L in lats]
    b.step(3, active=[True, False, True])
    torch.cuda.synchronize()
    assert torch.equal(lats[1].fields(), before[1])
    assert not torch.equal(lats[0].fields(), before[0])
    par, tim = b.pull_state()
    assert tim == [43, 40, 43]
    assert par == [L.cur for L in lats]


def test_graph_replay_equals_eager():
    nsteps = 2
    _, lats_a, ba, _ = _setup(T=16)
    _, lats_b, bb, _ = _setup(T=16)
    hg = ba.history_buffer(nsteps)
    ba.capture(nsteps, force_history=hg)
    for rep in range(3):
        ba.replay()
        he = bb.history_buffer(nsteps)
        bb.step(nsteps, force_history=he)
        torch.cuda.synchronize()
        assert torch.equal(hg, he), f"replay {rep} force history"
        for La, Lb in zip(lats_a, lats_b):
            assert torch.equal(La.fields(), Lb.fields()), f"replay {rep} state"
        # eager work between replays
        _ = lats_a[0].fields().sum().item()
    assert ba.pull_state() == bb.pull_state()


def test_history_layout_and_static_bodies():
    """A static mask (no capsule table) with a pose table: forces equal
    Lattice.forces on the same state."""
    e = _eels()[0]
    L = _lattice(e, 0)
    caps, poses = e.tables(10, 1)
    L.set_poses(poses[0])
    L.voxelize(caps[0])
    b = Batch([L])
    h = b.history_buffer(1)
    ref = L.forces().clone()
    b.step(1, force_history=h)
    torch.cuda.synchronize()
    assert h.shape == (1, 1, MAX_SOLIDS, 6)
    torch.testing.assert_close(h[0, 0, :EEL_SEGMENTS], ref, rtol=1e-12, atol=1e-15)


def test_batched_drag_within_one_percent_of_fp64():
    """Mean drag of batched FP32 eels against each eel alone in FP64 (the
    gather kernel, same kinematics): within the 1% bar (a reduced set: 6 eels
    of the batch workload, 300 steps)."""
    from openhomelbm_amd import inputs
    r = inputs.batched_drag_check(n=6, seed=3003, steps=300)
    print(r)
    assert r["mean_drag_rel_err_max"] < 0.01
    assert r["batch_mean_drag_rel_err"] < 0.01


def test_strip_heights_change_speed_only():
    """The same batch stepped with other strip heights (hl_batch_set_strips):
    states bitwise equal, force histories to summation-order rounding; the
    batch tuner leaves the state where it found it."""
    from openhomelbm_amd import autotune
    nsteps = 5
    _, lats_a, ba, _ = _setup()
    _, lats_b, bb, _ = _setup()
    bb.set_strips(8, 3)
    ha, hb = ba.history_buffer(nsteps), bb.history_buffer(nsteps)
    ba.step(nsteps, force_history=ha)
    bb.step(nsteps, force_history=hb)
    torch.cuda.synchronize()
    for La, Lb in zip(lats_a, lats_b):
        assert torch.equal(La.fields(), Lb.fields())
    scale = ha.abs().max().item()
    assert scale > 0
    torch.testing.assert_close(hb, ha, rtol=1e-10, atol=1e-12 * scale)
    before = [L.fields().clone() for L in lats_b]
    state = bb.pull_state()
    r = autotune.tune_batch(bb, steps=3, reps=1, shapes=((16, 4), (8, 3)))
    assert set(r["ms_per_step"]) == {"16x4", "8x3"} and tuple(r["strips"]) in ((16, 4), (8, 3))
    assert bb.pull_state() == state
    for L, f in zip(lats_b, before):
        assert torch.equal(L.fields(), f)
    with pytest.raises(Exception):
        bb.set_strips(2, 4)

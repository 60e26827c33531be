"""GPU parity of the persistent batched engine (hl_batch_run: every
environment advanced K steps in one cooperative launch, the wavefronts dealing
out the step strips, voxelization fused into the body strips' row loads, a
barrier per step).

Checks, on the eel batch of test_gpu_batch (uneven sizes, moving 8-capsule
eels re-voxelized every step):
* states equal the three-kernel schedule (hl_batch_step) and the FP32 oracle
  bit for bit; the force history equals hl_batch_step's bit for bit and the
  oracle's FP64 sums to summation-order rounding;
* one launch for all environments equals one launch per environment, states
  and force history bit for bit;
* the mask left behind is the last step's; inactive environments are
  untouched; no grid-barrier wait gives up.
"""
import numpy as np
import pytest
import torch

from openhomelbm_amd.inputs import EEL_SEGMENTS
from tests.test_gpu_batch import _oracle, _setup
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nsteps", [1, 6])
def test_run_matches_step_and_oracle(nsteps):
    t0 = 40
    eels, lats_a, ba, tabs = _setup(t0)
    _, lats_b, bb, _ = _setup(t0)
    m0 = [L.get_state(np.float32) for L in lats_a]
    ha, hb = ba.history_buffer(nsteps), bb.history_buffer(nsteps)
    ba.run(nsteps, force_history=ha)
    bb.step(nsteps, force_history=hb)
    torch.cuda.synchronize()
    assert not ba.run_timed_out()
    assert ba.pull_state() == bb.pull_state()
    hist_a, hist_b = ha.cpu().numpy(), hb.cpu().numpy()
    for i, (e, La, Lb) in enumerate(zip(eels, lats_a, lats_b)):
        assert torch.equal(La.fields(), Lb.fields()), f"env {i}: run differs from step"
        caps, poses, box = tabs[i]
        ref, ref_hist = _oracle(e, m0[i], caps, poses, box, t0, nsteps)
        assert np.array_equal(La.get_state(np.float32), ref), f"env {i}: run differs from the oracle"
        scale = np.abs(ref_hist).max()
        np.testing.assert_allclose(hist_a[:, i, :EEL_SEGMENTS], ref_hist, rtol=1e-10, atol=1e-12 * scale)
        assert np.array_equal(hist_a, hist_b)
        assert not hist_a[:, i, EEL_SEGMENTS:].any()
        assert torch.equal(La.mask, Lb.mask)


def test_run_all_equals_per_env():
    nsteps = 9
    _, lats_a, ba, _ = _setup()
    _, lats_b, bb, _ = _setup()
    ha, hb = ba.history_buffer(nsteps), bb.history_buffer(nsteps)
    ba.run(nsteps, force_history=ha)
    bb.run(nsteps, force_history=hb, per_env=True)
    torch.cuda.synchronize()
    assert not ba.run_timed_out() and not bb.run_timed_out()
    assert torch.equal(ha, hb)
    for La, Lb in zip(lats_a, lats_b):
        assert torch.equal(La.fields(), Lb.fields())


def test_run_leaves_last_mask_and_skips_inactive():
    t0 = 40
    eels, lats, b, tabs = _setup(t0)
    before = lats[1].fields().clone()
    b.run(5, active=[True, False, True])
    torch.cuda.synchronize()
    assert torch.equal(lats[1].fields(), before)
    par, tim = b.pull_state()
    assert tim == [t0 + 5, t0, t0 + 5] and par == [L.cur for L in lats]
    for i in (0, 2):
        e = eels[i]
        caps, _, box = tabs[i]
        want = np.zeros((e.ny, e.nx), np.uint8)
        orc.voxelize(orc.Env(9, e.nx, e.ny), caps[t0 + 4], box, want)
        assert np.array_equal(lats[i].mask.cpu().numpy().reshape(e.ny, e.nx), want)


def test_run_then_step_continue_alike():
    """Alternating the engines keeps parity, time and states consistent."""
    _, lats_a, ba, _ = _setup(T=24)
    _, lats_b, bb, _ = _setup(T=24)
    ba.run(3)
    ba.step(2)
    ba.run(4)
    bb.step(9)
    torch.cuda.synchronize()
    assert ba.pull_state() == bb.pull_state()
    for La, Lb in zip(lats_a, lats_b):
        assert torch.equal(La.fields(), Lb.fields())

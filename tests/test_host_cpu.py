"""Host-side logic that needs no GPU: the Strouhal estimator, the CPG
kinematics tables, the serving loop's kinematics windows and footprints, the
autotuner's size classes and the array handles of the solver API."""
import numpy as np
import pytest

from openhomelbm_amd import autotune, inputs, karman
from openhomelbm_amd.arrays import HostArray


def test_strouhal_recovers_synthetic_sinusoid():
    rng = np.random.default_rng(3)
    D, U = 16.0, 0.05
    for st in (0.16, 0.2137, 0.3):
        f = st * U / D
        t = np.arange(25000)
        lift = 0.3 + 0.05 * np.sin(2 * np.pi * f * t + 0.4) + 1e-3 * rng.standard_normal(len(t))
        assert abs(karman.strouhal(lift, D, U) - st) / st < 2e-3


def test_strouhal_picks_dominant_tone():
    t = np.arange(20000)
    lift = np.sin(2 * np.pi * 0.0005 * t) + 0.3 * np.sin(2 * np.pi * 0.002 * t)
    assert abs(karman.strouhal(lift, 1.0, 1.0) - 0.0005) < 5e-6


def test_eel_chain_geometry_and_speed_cap():
    for e in inputs.eel_batch(8, seed=11):
        caps, poses = e.tables(0, 50)
        assert caps.shape == (50, inputs.EEL_SEGMENTS, 8) and poses.shape == (50, inputs.EEL_SEGMENTS, 8)
        seg = np.linalg.norm(caps[..., 3:5] - caps[..., 0:2], axis=-1)
        np.testing.assert_allclose(seg, e.seg_len, rtol=1e-5)       # rigid segments
        np.testing.assert_allclose(caps[:, 1:, 0:2], caps[:, :-1, 3:5], atol=1e-4)  # joints connect
        np.testing.assert_allclose(poses[..., 2] ** 2 + poses[..., 3] ** 2, 1.0, rtol=1e-6)
        np.testing.assert_allclose(poses[1:, :, 4:8], poses[:-1, :, 0:4], atol=1e-5)  # last = previous step
        assert e.max_speed() <= inputs.EEL_VMAX * (1 + 1e-4)  # time scaled to the cap (sampled estimate)
        x0, x1, y0, y1, _, _ = e.box(0, 50)
        assert x0 <= caps[..., [0, 3]].min() - e.radius and caps[..., [0, 3]].max() + e.radius <= x1
        assert y0 <= caps[..., [1, 4]].min() - e.radius and caps[..., [1, 4]].max() + e.radius <= y1


def test_eel3d_rotation_tables():
    e = inputs.eel3d(nx=96, ny=48, nz=48, length=40.0, radius=4.0)
    caps, poses = e.tables(5, 4)
    R = poses[..., 3:12].reshape(4, e.nseg, 3, 3)
    np.testing.assert_allclose(R @ np.swapaxes(R, -1, -2), np.broadcast_to(np.eye(3), R.shape), atol=1e-6)
    d = caps[..., 3:6] - caps[..., 0:3]
    np.testing.assert_allclose(d / e.seg_len, R[..., :, 0], atol=1e-5)  # body x axis along the segment


def test_kinematics_window_roll_matches_modular_index():
    from openhomelbm_amd.serving import EnvSpec

    e = inputs.eel_batch(1, seed=2)[0]
    H, t0 = 64, 150
    caps, _ = e.tables(t0, H)
    rolled = np.roll(caps, t0 % H, axis=0)
    for t in (t0, t0 + 1, t0 + H - 1):
        np.testing.assert_array_equal(rolled[t % H], caps[t - t0])
    spec = EnvSpec(e, 9, horizon=H)
    assert spec.hbm_bytes() >= e.nx * e.ny * 48


def test_size_classes():
    assert autotune.size_class(27, 256, 128, 128) == "d3q27:4^11"
    assert autotune.size_class(27, 128, 64, 64) == "d3q27:4^9"
    assert autotune.size_class(9, 512, 128) == "d2q9:4^8"
    t = autotune.load_table()
    assert tuple(t["d3q27:4^11"]) in set(autotune.candidates())


def test_host_array_behaves_like_warp_array():
    changed = []
    a = HostArray(np.zeros((3, 2)), on_change=lambda: changed.append(1))
    a[1] = (4.0, 5.0)
    assert a.numpy()[1].tolist() == [4.0, 5.0] and changed
    a.assign(np.ones((3, 2)))
    assert np.all(a.numpy() == 1)
    a.zero_()
    assert not a.numpy().any()
    assert np.asarray(a).shape == (3, 2)


def test_first_step_wall_velocity_is_the_reference_fallback():
    """A 3D body whose transforms are not initialized: the pose pair built by
    first_step_last_pose, read the way the kernels read it (wall_velocity_3d:
    q = R^T (h - p), then R q + p minus R_last q + p_last), gives the
    reference's u_p = linear_v + angle_v x (h - p) (lbm_func_3d.py:464-474)
    for random poses, velocities and wall points; zero velocities give a body
    at rest."""
    import numpy as np
    from openhomelbm_amd.lbm3d import first_step_last_pose, quat_to_matrix
    rng = np.random.default_rng(5)
    for _ in range(50):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        R = quat_to_matrix(q)
        p = rng.uniform(10, 50, 3)
        row = np.concatenate([p, R.reshape(9)])
        v, w = rng.uniform(-0.05, 0.05, 3), rng.uniform(-0.01, 0.01, 3)
        last = first_step_last_pose(row, v, w)
        h = p + rng.uniform(-8, 8, 3)
        M, ML = row[3:12].reshape(3, 3), last[3:12].reshape(3, 3)
        qb = M.T @ (h - p)
        up = (M @ qb + p) - (ML @ qb + last[0:3])
        np.testing.assert_allclose(up, v + np.cross(w, h - p), rtol=0, atol=1e-12)
        rest = first_step_last_pose(row, np.zeros(3), np.zeros(3))
        assert np.array_equal(rest, row)

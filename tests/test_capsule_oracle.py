"""The capsule cut-link step of the oracle (hlo_capsule_cut and
hlo_step_capsules: the reference's ray query get_cutcell_multi_3d,
envs/lbm3d/lbm_func_3d.py:254-316, with closed-form capsule surfaces, and its
cut branch, lbm_func_3d.py:424-511) pinned independently of its formulas:

* crossings: hand-derived closed forms (axis and diagonal rays on spheres,
  capsule cylinders and caps, rays leaving from inside, rays parallel to the
  axis, misses, ties between solids, ignored solid ids), and a brute-force
  march in FP64 along every link of random capsules (the distance to the
  capsule's segment, sampled and bisected), which shares no formula with the
  oracle's quadratic intervals;
* a resting fluid of varying density around static bodies: every cut link
  pushes -2 w_i rho c_i, so the per-solid forces are closed forms of the
  brute-force cut set, and the torques sums of (hit - position) x F at the
  brute-force hit points;
* a sphere spinning on a rest state: forces from an independent NumPy
  evaluation (np_oracle's reconstruction and wall velocity) at the
  brute-force hit points; the same evaluation at the link midpoints is far
  off, so the hit point is what the oracle uses;
* a static body on a random state: every cell whose cut links (direction and
  solid) equal its mask's gives the mask rule's step bit for bit; a
  translating one within rounding; the cells where the sets differ do differ;
* the OpenMP build equals the serial one bit for bit.
"""
import numpy as np
import pytest

from oracle import np_oracle as npo
from oracle import oracle as orc
from tests.helpers import random_state

C3 = npo.D3_C
W32 = npo.D3_W.astype(np.float32)


def _dir(c):
    return int(np.where((C3 == np.array(c)).all(axis=1))[0][0])


def _poses(centers, rot_now=None, last_shift=None):
    P = np.zeros((len(centers), 24), np.float32)
    for k, p in enumerate(centers):
        R = np.eye(3) if rot_now is None or rot_now[k] is None else rot_now[k]
        P[k, 0:3] = p
        P[k, 3:12] = R.ravel()
        P[k, 12:15] = np.asarray(p) - (0 if last_shift is None else np.asarray(last_shift[k]))
        P[k, 15:24] = np.eye(3).ravel()
    return P


def _env(nx, ny, nz, poses, nu=0.02):
    env = orc.Env(27, nx, ny, nz, nu=nu, bc_type=(0, 1, 1, 1, 1, 1),
                  bc_value=((0.04, 0.0, 0.0),) + ((0.0, 0.0, 0.0),) * 5)
    env.solids = np.ascontiguousarray(poses, np.float32)
    return env


# ------------------------------------------------------------ brute force

def _seg_dist(P, a, b):
    """Distance of points P[n, 3] to segment ab (FP64)."""
    ab = b - a
    L2 = ab @ ab
    s = np.zeros(len(P)) if L2 == 0 else np.clip((P - a) @ ab / L2, 0.0, 1.0)
    return np.linalg.norm(P - (a + s[:, None] * ab), axis=1)


def brute_cuts(caps, nsolid, X, dirs, samples=4097):
    """Per direction c of dirs[n, 3]: (solid, t, margin) of the crossing
    nearest to X along X - c t, t in [0, 1] (solid -1, t -1: none), found by
    marching each capsule's signed distance (<= 0 inside) and bisecting its
    first sign change; margin = the smallest |distance - r| seen along the
    link at its ends and at local minima (small: grazing or touching, which
    the march cannot settle)."""
    X = np.asarray(X, np.float64)
    c = np.asarray(dirs, np.float64).reshape(-1, 3)
    n = len(c)
    ts = np.linspace(0.0, 1.0, samples)
    best = np.full(n, np.inf)
    sid = np.full(n, -1)
    margin = np.full(n, np.inf)
    for cap in np.asarray(caps, np.float64):
        s = int(cap[7])
        if s < 0 or s >= nsolid:
            continue
        a, b, r = cap[0:3], cap[3:6], cap[6]
        pts = X[None, None, :] - ts[None, :, None] * c[:, None, :]
        v = (_seg_dist(pts.reshape(-1, 3), a, b) - r).reshape(n, samples)
        # tangency shows as an end on the surface or a local minimum of the
        # distance near r; a transversal crossing is neither
        locmin = np.full(n, np.inf)
        dip = (v[:, 1:-1] <= v[:, :-2]) & (v[:, 1:-1] <= v[:, 2:])
        locmin = np.where(dip.any(axis=1), np.where(dip, np.abs(v[:, 1:-1]), np.inf).min(axis=1), np.inf)
        margin = np.minimum(margin, np.minimum(np.minimum(np.abs(v[:, 0]), np.abs(v[:, -1])), locmin))
        inside0 = v[:, 0] <= 0
        flip = (v <= 0) != inside0[:, None]
        has = flip.any(axis=1)
        if not has.any():
            continue
        k = np.nonzero(has)[0]
        j = flip[k].argmax(axis=1)
        lo, hi = ts[j - 1], ts[j]
        for _ in range(70):
            mid = 0.5 * (lo + hi)
            inside = _seg_dist(X[None, :] - mid[:, None] * c[k], a, b) - r <= 0
            same = inside == inside0[k]
            lo = np.where(same, mid, lo)
            hi = np.where(same, hi, mid)
        t = 0.5 * (lo + hi)
        better = t < best[k]
        best[k[better]] = t[better]
        sid[k[better]] = s
    return sid, np.where(sid >= 0, best, -1.0), margin


def brute_cut(caps, nsolid, X, c, samples=4097):
    sid, t, margin = brute_cuts(caps, nsolid, X, [c], samples)
    return int(sid[0]), float(t[0]), float(margin[0])


# ------------------------------------------------------------ crossings

SPHERE = np.array([[10, 10, 10, 10, 10, 10, 3, 0]], np.float32)
CAPSULE = np.array([[5, 10, 10, 15, 10, 10, 2, 0]], np.float32)


@pytest.mark.parametrize("precision,tol", [("f32", 2e-6), ("f64", 1e-12)])
@pytest.mark.parametrize("caps,cell,c,want", [
    (SPHERE, (13.25, 10, 10), (1, 0, 0), 0.25),                 # entering along -x
    (SPHERE, (13.25, 10, 10), (-1, 0, 0), None),                # pointing away
    (SPHERE, (13.5, 10, 10), (0, 1, 0), None),                  # passes beside
    (SPHERE, (10, 10, 10), (1, 0, 0), None),                    # from the centre: leaves at t = 3
    (SPHERE, (12, 12, 10), (-1, -1, 0), 3 / np.sqrt(2) - 2),    # from inside, leaving diagonally
    (SPHERE, (12.5, 12.5, 10), (1, 1, 0), 2.5 - 3 / np.sqrt(2)),  # entering diagonally
    (SPHERE, (12, 12, 12), (1, 1, 1), 2 - np.sqrt(3)),          # entering along a body diagonal
    (CAPSULE, (10, 12.5, 10), (0, 1, 0), 0.5),                  # cylinder from outside
    (CAPSULE, (10, 11, 10), (0, -1, 0), 1.0),                   # cylinder from inside, at t = 1
    (CAPSULE, (17.5, 10, 10), (1, 0, 0), 0.5),                  # end cap b
    (CAPSULE, (2.5, 10, 10), (-1, 0, 0), 0.5),                  # end cap a
    (CAPSULE, (14.5, 11, 10), (-1, 0, 0), None),                # along the axis inside: leaves at t > 1
    (CAPSULE, (16.2, 11, 10), (-1, 0, 0), 15 + np.sqrt(3) - 16.2),  # inside cap b, leaving along +x
    (CAPSULE, (10, 12.5, 10), (1, 0, 0), None),                 # parallel to the axis, outside
    # oblique onto the cylinder: (2.2 - t)^2 + (1 - t)^2 = 4
    (CAPSULE, (11, 12.2, 11), (0, 1, 1), (6.4 - np.sqrt(6.4 ** 2 - 8 * 1.84)) / 4),
])
def test_crossing_closed_forms(precision, tol, caps, cell, c, want):
    env = _env(24, 24, 24, _poses([(10, 10, 10)]))
    sid, t = orc.capsule_cut(env, caps, cell, _dir(c), precision)
    if want is None:
        assert sid == -1
    else:
        assert sid == 0 and abs(t - want) < tol, (t, want)


def test_crossing_ties_and_ignored_solids():
    env = _env(24, 24, 24, _poses([(10, 10, 10), (10, 10, 10)]))
    twins = np.array([[10, 10, 10, 10, 10, 10, 3, 1], [10, 10, 10, 10, 10, 10, 3, 0]], np.float32)
    assert orc.capsule_cut(env, twins, (13.25, 10, 10), _dir((1, 0, 0)))[0] == 1  # first capsule wins a tie
    near = np.array([[10, 10, 10, 10, 10, 10, 3, 1], [10, 10, 10, 10, 10, 10, 3.1, 0]], np.float32)
    sid, t = orc.capsule_cut(env, near, (13.25, 10, 10), _dir((1, 0, 0)))
    assert sid == 0 and abs(t - 0.15) < 2e-6  # the larger sphere is met first
    beyond = np.array([[10, 10, 10, 10, 10, 10, 3, 2]], np.float32)  # solid id >= nsolid
    assert orc.capsule_cut(env, beyond, (13.25, 10, 10), _dir((1, 0, 0)))[0] == -1
    assert orc.capsule_cut(env, SPHERE, (13.25, 10, 10), 0)[0] == -1  # the rest direction is never cut


def test_crossings_against_brute_force():
    rng = np.random.default_rng(5)
    caps = []
    for k in range(12):
        a = rng.uniform(8, 22, 3)
        d = rng.standard_normal(3)
        d *= rng.uniform(0, 7) / np.linalg.norm(d)
        caps.append(list(a) + list(a + d if k % 4 else a) + [rng.uniform(1.2, 4.0), k % 3])
    caps = np.array(caps, np.float32)
    env = _env(32, 32, 32, _poses([(15, 15, 15)] * 3))
    checked = hits = 0
    for k in range(len(caps)):
        one = caps[k:k + 1]
        mid = 0.5 * (caps[k, 0:3] + caps[k, 3:6])
        for _ in range(14):
            X = mid + rng.uniform(-1, 1, 3) * (np.abs(caps[k, 3:6] - caps[k, 0:3]) / 2 + caps[k, 6] + 2)
            X = X.astype(np.float32).astype(np.float64)
            for i in range(1, 27):
                sid_b, t_b, margin = brute_cut(one, 3, X, C3[i])
                if margin < 1e-5:
                    continue  # grazing or touching: the march cannot settle it
                sid64, t64 = orc.capsule_cut(env, one, X, i, "f64")
                sid32, t32 = orc.capsule_cut(env, one, X, i, "f32")
                assert sid64 == sid_b, (k, X, i, sid64, sid_b, t64, t_b)
                checked += 1
                if sid_b >= 0:
                    hits += 1
                    assert abs(t64 - t_b) < 1e-9, (k, X, i, t64, t_b)
                    if abs(t_b) > 1e-4 and abs(t_b - 1) > 1e-4:
                        assert sid32 == sid_b and abs(t32 - t_b) < 1e-4, (k, X, i, t32, t_b)
    # the nearest crossing over all capsules at once
    for _ in range(300):
        X = rng.uniform(6, 24, 3).astype(np.float32).astype(np.float64)
        i = int(rng.integers(1, 27))
        sid_b, t_b, margin = brute_cut(caps, 3, X, C3[i])
        if margin < 1e-5:
            continue
        sid64, t64 = orc.capsule_cut(env, caps, X, i, "f64")
        if sid_b >= 0:
            assert abs(t64 - t_b) < 1e-9
    assert checked > 3000 and hits > 200, (checked, hits)


# ------------------------------------------------------------ the step

def _scene():
    """A sphere and an oblique capsule (solids 0 and 1) in a 40x36x34 box."""
    caps = np.array([[13.3, 17.6, 16.2, 13.3, 17.6, 16.2, 4.6, 0],
                     [24.1, 15.2, 14.7, 31.4, 20.9, 19.3, 3.1, 1]], np.float32)
    return caps, [(13.3, 17.6, 16.2), (27.75, 18.05, 17.0)]


def _interior_links(env, caps, region):
    """Brute-force cut links {(z, y, x, i): (solid, t)} of the interior
    cells in region (cells farther than 2 from every surface have none)."""
    out = {}
    nsolid = len(env.solids)
    capd = np.asarray(caps, np.float64)
    (x0, x1), (y0, y1), (z0, z1) = region
    for z in range(max(z0, 1), min(z1, env.nz - 1)):
        for y in range(max(y0, 1), min(y1, env.ny - 1)):
            for x in range(max(x0, 1), min(x1, env.nx - 1)):
                X = np.array([[x, y, z]], np.float64)
                if min(abs(_seg_dist(X, c[0:3], c[3:6])[0] - c[6]) for c in capd) > 2.0:
                    continue
                sid, t, margin = brute_cuts(caps, nsolid, X[0], C3[1:], samples=513)
                assert margin.min() > 1e-6, "scene has a grazing link"
                for i in np.nonzero(sid >= 0)[0]:
                    out[(z, y, x, int(i) + 1)] = (int(sid[i]), float(t[i]))
    return out


REGION = ((5, 38), (8, 28), (8, 26))


@pytest.mark.slow
def test_resting_fluid_forces_are_closed_forms():
    """u = 0, S = 0 and a density that varies with position: a cut link of a
    cell of density rho pushes -2 fl(w_i rho) c_i (u_p = 0), so the forces are
    closed forms of the brute-force cut set and the torques sums at its hit
    points. (With a uniform density the links on either side of a crossing
    cancel: every interior cell takes the update, inside a body or not.)"""
    caps, centers = _scene()
    env = _env(40, 36, 34, _poses(centers))
    z, y, x = np.meshgrid(*[np.arange(n) for n in env.shape], indexing="ij")
    m0 = np.zeros((10,) + env.shape, np.float32)
    m0[0] = 1.0 + 0.004 * x - 0.003 * y + 0.002 * z
    _, F = orc.step_capsules(env, m0, caps)
    links = _interior_links(env, caps, REGION)
    want = np.zeros((2, 6))
    for (z, y, x, i), (sid, t) in links.items():
        f = -2.0 * np.float64(W32[i] * m0[0, z, y, x]) * C3[i]
        h = np.array([x, y, z], np.float64) - C3[i] * t
        want[sid, 0:3] += f
        want[sid, 3:6] += np.cross(h - np.array(centers[sid]), f)
    assert len(links) > 1500
    np.testing.assert_allclose(F[:, 0:3], want[:, 0:3], rtol=0, atol=1e-12)
    np.testing.assert_allclose(F[:, 3:6], want[:, 3:6], rtol=0, atol=2e-5)
    assert np.abs(want[:, 3:6]).max() > 1e-2  # the hit points carry a torque the check sees


def _link_forces_np(env, m, links, at_mid):
    """Per-solid forces of the cut links, rebuilt in FP64 with np_oracle's
    reconstruction and wall velocity at the hit point (or link midpoint)."""
    inv = npo.inverse_index(C3)
    out = np.zeros((len(env.solids), 6))
    for (z, y, x, i), (sid, t) in links.items():
        rho, u, S = npo._unpack(m[:, z, y, x].astype(np.float64), 27)
        h = np.array([x, y, z], np.float64) - C3[i] * (0.5 if at_mid else t)
        up, r = npo._wall_velocity(27, env.solids, np.array([sid]), h[:, None])
        up, r = up[:, 0], r[:, 0]
        Ss = np.outer(up, up) + S - np.outer(u, u)
        pop = npo.recon(27, rho, up, Ss, i)
        cur = npo.recon(27, rho, u, S, inv[i])
        f = cur * (C3[inv[i]] - up) - pop * (C3[i] - up)
        out[sid, 0:3] += f
        out[sid, 3:6] += np.cross(r, f)
    return out


@pytest.mark.slow
def test_spinning_sphere_forces_use_the_hit_point():
    caps = np.array([[16.3, 15.7, 14.2, 16.3, 15.7, 14.2, 5.4, 0]], np.float32)
    ang = 0.02
    Rz = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
    env = _env(32, 32, 30, _poses([(16.3, 15.7, 14.2)], rot_now=[Rz]))
    m0 = orc.init_uniform(env, (0.0, 0.0, 0.0))
    _, F = orc.step_capsules(env, m0, caps)
    links = _interior_links(env, caps, ((8, 25), (7, 25), (6, 23)))
    at_hit = _link_forces_np(env, m0, links, at_mid=False)
    at_mid = _link_forces_np(env, m0, links, at_mid=True)
    # FP32 against FP64: within 5e-5 of torques near 68; the midpoints move
    # the result forty times farther
    assert np.abs(at_hit).max() > 10
    np.testing.assert_allclose(F, at_hit, rtol=0, atol=5e-5)
    assert np.abs(at_mid - at_hit).max() > 40 * 5e-5


def _mask_links(env, mask):
    """Mask-rule cut links {(z, y, x, i): solid} of the interior cells."""
    occ = mask != 0
    out = {}
    zs, ys, xs = np.nonzero(occ)
    for z in range(1, env.nz - 1):
        for y in range(1, env.ny - 1):
            for x in range(1, env.nx - 1):
                for i in range(1, 27):
                    cx, cy, cz = C3[i]
                    s = (z - cz, y - cy, x - cx)
                    if occ[z, y, x] != occ[s]:
                        out[(z, y, x, i)] = int(mask[z, y, x] if occ[z, y, x] else mask[s]) - 1
    return out


@pytest.mark.slow
@pytest.mark.parametrize("moving", [False, True])
def test_cells_with_the_mask_cut_set_take_the_mask_step(moving):
    caps, centers = _scene()
    shift = [(0.012, -0.004, 0.003), (-0.006, 0.01, 0.0)] if moving else None
    env = _env(40, 36, 34, _poses(centers, last_shift=shift))
    env.mask = orc.voxelize(env, caps)
    m0 = random_state(27, env.shape, seed=61).astype(np.float32)
    got, _ = orc.step_capsules(env, m0, caps)
    ref = orc.step(env, m0, "f32")
    mlinks = _mask_links(env, env.mask)
    clinks = {}
    for z in range(1, env.nz - 1):
        for y in range(1, env.ny - 1):
            for x in range(1, env.nx - 1):
                if not (REGION[2][0] <= z < REGION[2][1] and REGION[1][0] <= y < REGION[1][1]
                        and REGION[0][0] <= x < REGION[0][1]):
                    continue
                for i in range(1, 27):
                    sid, _t = orc.capsule_cut(env, caps, (x, y, z), i)
                    if sid >= 0:
                        clinks[(z, y, x, i)] = sid
    cells = {k[:3] for k in list(mlinks) + list(clinks)}
    same = [c for c in cells if all(mlinks.get(c + (i,)) == clinks.get(c + (i,)) for i in range(1, 27))]
    differ = sorted(cells - set(same))
    assert len(same) > 300 and len(differ) > 20, (len(same), len(differ))
    sel = tuple(np.array(same).T)
    if moving:
        np.testing.assert_allclose(got[(slice(None),) + sel], ref[(slice(None),) + sel], rtol=0, atol=2e-6)
    else:
        assert np.array_equal(got[(slice(None),) + sel], ref[(slice(None),) + sel])
    dsel = tuple(np.array(differ).T)
    assert (np.abs(got[(slice(None),) + dsel] - ref[(slice(None),) + dsel]).max(axis=0) > 0).mean() > 0.9
    # away from every link of either set the step is the plain step
    far = np.ones(env.shape, bool)
    far[tuple(np.array(sorted(cells)).T)] = False
    far[[0, -1], :, :] = far[:, [0, -1], :] = far[:, :, [0, -1]] = False
    assert np.array_equal(got[:, far], ref[:, far])


def test_openmp_capsule_step_is_bitwise_serial():
    caps, centers = _scene()
    env = _env(40, 36, 34, _poses(centers, rot_now=[None, np.array([[1, 0, 0], [0, 0.9998, -0.02], [0, 0.02, 0.9998]])]))
    m0 = random_state(27, env.shape, seed=62).astype(np.float32)
    a, _ = orc.step_capsules(env, m0, caps)
    b, _ = orc.step_capsules(env, m0, caps, omp=True)
    assert np.array_equal(a, b)
    d, _ = orc.step_capsules(env, m0.astype(np.float64), caps, "f64")
    assert np.abs(a - d).max() < 1e-5

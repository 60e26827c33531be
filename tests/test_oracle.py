"""Pins for the CPU oracle: closed forms, moment identities, invariants, and
agreement of the two independent restatements (C and NumPy).

The reference ships no golden vectors and cannot run here (Warp is not
installed), so these are the checks that pin the restatement."""
import numpy as np
import pytest

from oracle import np_oracle as npo
from oracle import oracle as orc
from tests.helpers import random_state, tensor_rel_l2


# ------------------------------------------------------------- lattices

@pytest.mark.parametrize("lattice", [9, 27])
def test_lattice_isotropy(lattice):
    C = (npo.D2_C if lattice == 9 else npo.D3_C).astype(float)
    W = npo.D2_W if lattice == 9 else npo.D3_W
    d = C.shape[1]
    assert abs(W.sum() - 1) < 1e-15
    assert np.allclose(W @ C, 0, atol=1e-15)
    assert np.allclose(np.einsum("i,ia,ib->ab", W, C, C), np.eye(d) / 3, atol=1e-15)
    m4 = np.einsum("i,ia,ib,ic,id->abcd", W, C, C, C, C)
    I = np.eye(d)
    iso = (np.einsum("ab,cd->abcd", I, I) + np.einsum("ac,bd->abcd", I, I) + np.einsum("ad,bc->abcd", I, I)) / 9
    assert np.allclose(m4, iso, atol=1e-15)
    inv = npo.inverse_index(C.astype(int))
    expect = [0, 2, 1, 4, 3, 7, 8, 5, 6] if lattice == 9 else \
        [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17, 20, 19, 22, 21, 24, 23, 26, 25]
    assert list(inv) == expect


# -------------------------------------------------------- reconstruction

def _recon_all(lattice, m, precision="f64"):
    q = 9 if lattice == 9 else 27
    return np.array([orc.recon(lattice, m, i, precision) for i in range(q)])


@pytest.mark.parametrize("seed", range(5))
def test_d2q9_recon_moments(seed):
    """Sum f = rho, sum c f = rho u, sum cc f = rho (I/3 + S), and the third
    moment sum cx^2 cy f = rho (uy/3 + A_xxy) (lbm_func.py:695-696)."""
    rng = np.random.default_rng(seed)
    rho = 1 + 0.1 * rng.standard_normal()
    ux, uy = 0.1 * rng.standard_normal(2)
    sxx, syy, sxy = 0.01 * rng.standard_normal(3)
    f = _recon_all(9, [rho, ux, uy, sxx, syy, sxy])
    C = npo.D2_C.astype(float)
    assert abs(f.sum() - rho) < 1e-14
    assert np.allclose(C.T @ f, rho * np.array([ux, uy]), atol=1e-14)
    Pi = np.einsum("ia,ib,i->ab", C, C, f)
    assert np.allclose(Pi, rho * (np.eye(2) / 3 + np.array([[sxx, sxy], [sxy, syy]])), atol=1e-14)
    a_xxy = sxx * uy + 2 * sxy * ux - 2 * ux * ux * uy
    a_xyy = syy * ux + 2 * sxy * uy - 2 * ux * uy * uy
    assert abs(np.sum(C[:, 0] ** 2 * C[:, 1] * f) - rho * (uy / 3 + a_xxy)) < 1e-14
    assert abs(np.sum(C[:, 0] * C[:, 1] ** 2 * f) - rho * (ux / 3 + a_xyy)) < 1e-14
    # the NumPy restatement rebuilds the same populations
    rho_, u_, S_ = npo._unpack(np.array([rho, ux, uy, sxx, syy, sxy]), 9)
    g = np.array([npo.recon(9, rho_, u_, S_, i) for i in range(9)])
    assert np.allclose(f, g, atol=1e-15)


@pytest.mark.parametrize("seed", range(5))
def test_d3q27_recon_moments(seed):
    """Sum f = rho, sum c f = rho u, sum cc f = rho (I/3 + uu + S/3)
    (the literal second moment of lbm_func_3d.py:224-251)."""
    rng = np.random.default_rng(100 + seed)
    rho = 1 + 0.1 * rng.standard_normal()
    u = 0.1 * rng.standard_normal(3)
    s = 0.01 * rng.standard_normal(6)  # xx, yy, zz, xy, xz, yz
    m = np.concatenate([[rho], u, s])
    f = _recon_all(27, m)
    C = npo.D3_C.astype(float)
    S = np.array([[s[0], s[3], s[4]], [s[3], s[1], s[5]], [s[4], s[5], s[2]]])
    assert abs(f.sum() - rho) < 1e-14
    assert np.allclose(C.T @ f, rho * u, atol=1e-14)
    Pi = np.einsum("ia,ib,i->ab", C, C, f)
    assert np.allclose(Pi, rho * (np.eye(3) / 3 + np.outer(u, u) + S / 3), atol=1e-14)
    rho_, u_, S_ = npo._unpack(m, 27)
    g = np.array([npo.recon(27, rho_, u_, S_, i) for i in range(27)])
    assert np.allclose(f, g, atol=1e-15)


@pytest.mark.parametrize("lattice", [9, 27])
def test_recon_f32_close_to_f64(lattice):
    rng = np.random.default_rng(7)
    n = 6 if lattice == 9 else 10
    m = np.concatenate([[1.0], 0.05 * rng.standard_normal(n - 1)])
    a = _recon_all(lattice, m, "f32")
    b = _recon_all(lattice, m, "f64")
    assert np.allclose(a, b, rtol=3e-6, atol=1e-8)


# -------------------------------------------------------------- init

@pytest.mark.parametrize("lattice", [9, 27])
def test_init_uniform_closed_form(lattice):
    """set_uniform_flow subtracts the raw pixx from every diagonal
    (karman_lbm_env.py:37-40): Sxx = 0, Syy = uy^2 - ux^2, Szz = uz^2 - ux^2,
    Sxy = ux uy (the equilibrium second moment is rho (I/3 + uu))."""
    u = (0.05, -0.02, 0.03)
    env = orc.Env(lattice, 7, 5, 4 if lattice == 27 else 1)
    m = orc.init_uniform(env, u, "f64")
    if lattice == 9:
        ux, uy = u[:2]
        expect = [1, ux, uy, 0, uy * uy - ux * ux, ux * uy]
    else:
        ux, uy, uz = u
        expect = [1, ux, uy, uz, 0, uy * uy - ux * ux, uz * uz - ux * ux, ux * uy, ux * uz, uy * uz]
    for f, v in enumerate(expect):
        assert np.allclose(m[f], v, atol=1e-15), f
    m32 = orc.init_uniform(env, u, "f32")
    assert m32.dtype == np.float32 and np.allclose(m32, m, atol=1e-7)


# -------------------------------------------------------------- step

def _masked_env(lattice, shape, seed, moving=True):
    """A small env with two solids (one moving) voxelized from capsules."""
    rng = np.random.default_rng(seed)
    if lattice == 9:
        ny, nx = shape
        env = orc.Env(9, nx, ny, nu=0.03, bc_type=(0, 1, 2, 1), bc_value=((0.04, 0.01, 0),) + ((0, 0, 0),) * 5)
        caps = np.array([[nx * 0.3, ny * 0.5, 0, nx * 0.3, ny * 0.5, 0, 2.2, 0],
                         [nx * 0.6, ny * 0.35, 0, nx * 0.75, ny * 0.6, 0, 1.6, 1]], np.float32)
        ang = 0.05 if moving else 0.0
        solids = np.array([
            [nx * 0.3, ny * 0.5, 1, 0, nx * 0.3, ny * 0.5, 1, 0],
            [nx * 0.67, ny * 0.47, np.cos(ang), np.sin(ang), nx * 0.67 - (0.02 if moving else 0),
             ny * 0.47 + (0.01 if moving else 0), 1, 0]], np.float32)
        force = 1e-4 * rng.standard_normal((2, ny, nx)).astype(np.float32)
    else:
        nz, ny, nx = shape
        env = orc.Env(27, nx, ny, nz, nu=0.02, bc_type=(0, 1, 2, 1, 1, 2),
                      bc_value=((0.04, 0.0, 0.01),) + ((0, 0, 0),) * 5)
        caps = np.array([[nx * 0.35, ny * 0.5, nz * 0.5, nx * 0.35, ny * 0.5, nz * 0.5, 1.8, 0],
                         [nx * 0.6, ny * 0.3, nz * 0.4, nx * 0.7, ny * 0.6, nz * 0.6, 1.3, 1]], np.float32)
        ang = 0.04 if moving else 0.0
        c, s = np.cos(ang), np.sin(ang)
        Rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
        I = np.eye(3)
        p1 = np.array([nx * 0.65, ny * 0.45, nz * 0.5])
        solids = np.zeros((2, 24), np.float32)
        solids[0, 0:3] = [nx * 0.35, ny * 0.5, nz * 0.5]
        solids[0, 3:12] = I.ravel()
        solids[0, 12:15] = solids[0, 0:3]
        solids[0, 15:24] = I.ravel()
        solids[1, 0:3] = p1
        solids[1, 3:12] = Rz.ravel()
        solids[1, 12:15] = p1 - (np.array([0.02, -0.01, 0.005]) if moving else 0)
        solids[1, 15:24] = I.ravel()
        force = 1e-4 * rng.standard_normal((3,) + tuple(shape)).astype(np.float32)
    mask = orc.voxelize(env, caps)
    env.mask = mask
    env.solids = solids
    env.force = force
    return env, caps


@pytest.mark.parametrize("lattice,shape", [(9, (10, 14)), (27, (7, 8, 11))])
@pytest.mark.parametrize("with_body", [False, True])
def test_c_oracle_matches_numpy_restatement(lattice, shape, with_body):
    """The C FP64 oracle and the independent NumPy restatement agree to
    rounding on seeded states: bodies (static and moving), body force,
    velocity / zero-gradient / free-slip faces."""
    if with_body:
        env, _ = _masked_env(lattice, shape, 3)
        assert env.mask.any()
    else:
        env = orc.Env(lattice, shape[-1], shape[-2], shape[0] if lattice == 27 else 1, nu=0.05,
                      bc_type=(0, 1, 1, 1, 1, 1), bc_value=((0.05, 0.0, 0.0),) + ((0, 0, 0),) * 5)
    m = random_state(lattice, shape, seed=11)
    a = orc.step(env, m, "f64")
    # the C interface carries nu and face values as float32
    nu32 = float(np.float32(env.nu))
    bcv32 = [[float(np.float32(v)) for v in row] for row in env.bc_value]
    b = npo.step(lattice, m, nu32, env.bc_type, bcv32, env.mask, env.force, env.solids)
    err = tensor_rel_l2(lattice, a, b)
    assert max(err.values()) < 1e-13, err
    if with_body:
        fa = orc.compute_forces(env, m, "f64")
        fb = npo.forces(lattice, m, env.mask, env.solids)
        assert np.allclose(fa, fb, rtol=1e-11, atol=1e-15), (fa, fb)
        assert np.abs(fa[:, :2]).max() > 1e-6  # forces are not trivially zero


@pytest.mark.parametrize("lattice,shape", [(9, (12, 16)), (27, (8, 9, 10))])
def test_f32_oracle_close_to_f64(lattice, shape):
    env, _ = _masked_env(lattice, shape, 5)
    m = random_state(lattice, shape, seed=12)
    a = orc.step(env, m.astype(np.float32), "f32")
    b = orc.step(env, m, "f64")
    err = tensor_rel_l2(lattice, a, b)
    assert err["rho"] < 1e-6 and err["u"] < 1e-5 and err["S"] < 1e-6, err
    c = orc.step(env, m.astype(np.float32), "f32", omp=True)
    assert np.array_equal(a, c)  # OpenMP build is the same computation


@pytest.mark.parametrize("lattice,shape", [(9, (16, 20)), (27, (10, 12, 14))])
@pytest.mark.parametrize("with_body", [False, True])
def test_rest_state_is_fixed_point(lattice, shape, with_body):
    """rho = 1, u = 0, S = 0 is a steady state of the step: every pulled
    population is w_i, every wall population is w_i, and the collision of the
    rest equilibrium returns it. Holds to rounding in FP64; in FP32 the
    float32-rounded D3Q27 weights do not sum to exactly 1, so rho drifts by at
    most a few ulp per step while the field stays uniform."""
    env = orc.Env(lattice, shape[-1], shape[-2], shape[0] if lattice == 27 else 1, nu=0.02)
    m = orc.init_uniform(env, (0, 0, 0), "f64")
    if with_body:
        env2, _ = _masked_env(lattice, shape, 1, moving=False)
        env.mask, env.solids = env2.mask, env2.solids
    for prec in ("f64", "f32"):
        cur = m.astype(np.float32 if prec == "f32" else np.float64)
        for _ in range(3):
            cur = orc.step(env, cur, prec)
        tol = 1e-15 if prec == "f64" else 3 * 3e-7
        assert np.abs(cur[0] - 1).max() < tol
        assert np.abs(cur[1:]).max() < tol
        assert np.all(cur[0] == cur[0].flat[0])
    if with_body:
        # bodies clear of the wall layer: every cut link pairs with its reverse
        occ = env.mask != 0
        assert not (occ[0].any() or occ[-1].any() or occ[..., 0].any() or occ[..., -1].any())
        f = orc.compute_forces(env, m, "f64")
        assert np.abs(f).max() < 1e-12  # rest fluid pushes on no body


def _uniform_interior_state(lattice, shape, m0):
    return np.broadcast_to(np.asarray(m0, dtype=np.float64)[(...,) + (None,) * len(shape)],
                           (len(m0),) + tuple(shape)).copy()


@pytest.mark.parametrize("omega", [0.7, 1.0, 1.6])
def test_d2q9_collision_closed_forms(omega):
    """On a uniform state the pull returns the cell's own populations, so an
    interior cell shows the bare collision: the trace of Pi relaxes to
    rho (2/3 + u.u) in one step, the deviator relaxes as
    dev' = (1 - Omega) dev + Omega dev(rho uu), and at Omega = 1 the result is
    the equilibrium S = uu."""
    nu = (1 / omega - 0.5) / 3
    rho, ux, uy = 1.02, 0.04, -0.03
    s = np.array([0.002, -0.001, 0.0015])
    m0 = [rho, ux, uy, *s]
    env = orc.Env(9, 6, 5, nu=nu)
    out = orc.step(env, _uniform_interior_state(9, (5, 6), m0), "f64")[:, 2, 3]
    assert abs(out[0] - rho) < 1e-14 and abs(out[1] - ux) < 1e-15 and abs(out[2] - uy) < 1e-15
    Pi = rho * (np.eye(2) / 3 + np.array([[s[0], s[2]], [s[2], s[1]]]))
    P2 = rho * (np.eye(2) / 3 + np.array([[out[3], out[5]], [out[5], out[4]]]))
    assert abs(np.trace(P2) - rho * (2 / 3 + ux * ux + uy * uy)) < 1e-14
    dev = lambda A: A - np.trace(A) / 2 * np.eye(2)
    ruu = rho * np.outer([ux, uy], [ux, uy])
    assert np.allclose(dev(P2), (1 - omega) * dev(Pi) + omega * dev(ruu), atol=1e-14)
    if omega == 1.0:
        assert np.allclose(out[3:], [ux * ux, uy * uy, ux * uy], atol=1e-14)


@pytest.mark.parametrize("omega", [0.8, 1.0, 1.7])
def test_d3q27_collision_closed_forms(omega):
    """As above in 3D, where the rebuilt second moment is rho (I/3 + uu + S/3):
    trace(Pi') = rho (1 + u.u), dev' = (1 - Omega) dev(Pi) + Omega dev(rho uu),
    and Omega = 1 gives S' = uu."""
    nu = (1 / omega - 0.5) / 3
    rho = 0.98
    u = np.array([0.03, -0.02, 0.01])
    s = np.array([0.002, -0.001, 0.0005, 0.0007, -0.0004, 0.0003])
    m0 = [rho, *u, *s]
    env = orc.Env(27, 5, 5, 5, nu=nu)
    out = orc.step(env, _uniform_interior_state(27, (5, 5, 5), m0), "f64")[:, 2, 2, 2]
    assert abs(out[0] - rho) < 1e-14 and np.allclose(out[1:4], u, atol=1e-15)
    S = np.array([[s[0], s[3], s[4]], [s[3], s[1], s[5]], [s[4], s[5], s[2]]])
    Pi = rho * (np.eye(3) / 3 + np.outer(u, u) + S / 3)
    S2 = np.array([[out[4], out[7], out[8]], [out[7], out[5], out[9]], [out[8], out[9], out[6]]])
    P2 = rho * (np.eye(3) / 3 + S2)
    assert abs(np.trace(P2) - rho * (1 + u @ u)) < 1e-14
    dev = lambda A: A - np.trace(A) / 3 * np.eye(3)
    assert np.allclose(dev(P2), (1 - omega) * dev(Pi) + omega * dev(rho * np.outer(u, u)), atol=1e-14)
    if omega == 1.0:
        assert np.allclose(S2, np.outer(u, u), atol=1e-14)


@pytest.mark.parametrize("lattice,shape", [(9, (8, 9)), (27, (6, 6, 7))])
def test_boundary_closed_form(lattice, shape):
    """Face cells: rho_b = rho_n, u_b = bc value (type 0) / u_n (type 1) /
    tangential u_n (type 2), and S_b = u_b u_b - u_n u_n + S_n, because the
    equilibrium S of f_eq is uu whenever no population is clamped
    (lbm_func.py:634-682, lbm_func_3d.py:627-686)."""
    d = 2 if lattice == 9 else 3
    env = orc.Env(lattice, shape[-1], shape[-2], shape[0] if d == 3 else 1, nu=0.04,
                  bc_type=(0, 2, 1, 1, 2, 1), bc_value=((0.03, 0.01, -0.02),) + ((0, 0, 0),) * 5)
    m = random_state(lattice, shape, seed=21)
    out = orc.step(env, m, "f64")
    rho, u, S = npo._unpack(out, lattice)
    # left face (x = 0), away from edges: type 0
    if d == 2:
        b, n = (slice(1, -1), 0), (slice(1, -1), 1)
    else:
        b, n = (slice(1, -1), slice(1, -1), 0), (slice(1, -1), slice(1, -1), 1)
    assert np.array_equal(rho[b], rho[n])
    v = np.array(env.bc_value[0][:d], dtype=np.float32).astype(np.float64)  # carried as float32
    for a in range(d):
        assert np.all(u[a][b] == v[a])
    for a in range(d):
        for c in range(d):
            expect = v[a] * v[c] - u[a][n] * u[c][n] + S[a, c][n]
            assert np.allclose(S[a, c][b], expect, atol=1e-15)
    # type 2 (free slip): D2Q9 face id 1 is the top (y = ny-1), D3Q27 face
    # id 1 the right (x = nx-1); the normal velocity is zero, the rest copied
    if d == 2:
        b, n = (-1, slice(1, -1)), (-2, slice(1, -1))
        assert np.all(u[1][b] == 0) and np.array_equal(u[0][b], u[0][n])
    else:
        b, n = (slice(1, -1), slice(1, -1), -1), (slice(1, -1), slice(1, -1), -2)
        assert np.all(u[0][b] == 0) and np.array_equal(u[1][b], u[1][n]) and np.array_equal(u[2][b], u[2][n])


# ------------------------------------------------------------ voxelizer

def test_voxelizer_disc_integer_closed_form():
    """Disc of diameter 16 at (96, 65) on 512x128: inside iff
    (x-96)^2 + (y-65)^2 <= 64, evaluated in integers."""
    env = orc.Env(9, 512, 128)
    mask = orc.voxelize(env, [96, 65, 0, 96, 65, 0, 8, 0])
    y, x = np.mgrid[0:128, 0:512]
    expect = ((x - 96) ** 2 + (y - 65) ** 2 <= 64).astype(np.uint8)
    assert np.array_equal(mask, expect)


def test_voxelizer_sphere_integer_closed_form():
    """Sphere of diameter 24 at (64, 64, 65): (x-64)^2+(y-64)^2+(z-65)^2 <= 144."""
    env = orc.Env(27, 128, 96, 112)
    mask = orc.voxelize(env, [64, 64, 65, 64, 64, 65, 12, 0], box=(40, 90, 40, 90, 40, 90))
    z, y, x = np.mgrid[0:112, 0:96, 0:128]
    expect = ((x - 64) ** 2 + (y - 64) ** 2 + (z - 65) ** 2 <= 144).astype(np.uint8)
    assert np.array_equal(mask, expect)


@pytest.mark.parametrize("seed", range(3))
def test_voxelizer_capsules_match_vectorised_and_bruteforce(seed):
    rng = np.random.default_rng(seed)
    shape = (20, 24, 28)
    caps = []
    for k in range(5):
        a = rng.uniform(4, 20, 3)
        b = a + rng.uniform(-6, 6, 3)
        caps.append([*a, *b, rng.uniform(1.0, 3.5), k])
    caps = np.array(caps, np.float32)
    env = orc.Env(27, shape[2], shape[1], shape[0])
    m = orc.voxelize(env, caps)
    assert np.array_equal(m, npo.voxelize_f32(shape, caps))  # bit-exact FP32
    # brute force in FP64, ignoring cells within 1e-4 of a surface
    z, y, x = np.mgrid[0:shape[0], 0:shape[1], 0:shape[2]]
    p = np.stack([x, y, z], -1).astype(float)
    inside_any = np.zeros(shape, bool)
    ambiguous = np.zeros(shape, bool)
    for c in caps.astype(float):
        a, b, r = c[0:3], c[3:6], c[6]
        ab = b - a
        t = np.clip(((p - a) @ ab) / (ab @ ab), 0, 1)
        dist = np.linalg.norm(p - (a + t[..., None] * ab), axis=-1)
        inside_any |= dist <= r
        ambiguous |= np.abs(dist - r) < 1e-4
    assert np.array_equal((m != 0)[~ambiguous], inside_any[~ambiguous])


# ------------------------------------------------------ wall rule / forces

def test_wall_velocity_translation_and_rotation():
    """u_p of a translating body is its displacement; of a body rotating by
    dtheta about p it is R(dtheta) q - q at the body-frame point q."""
    env = orc.Env(9, 40, 30)
    mask = orc.voxelize(env, [20, 15, 0, 20, 15, 0, 5, 0])
    m = orc.init_uniform(env, (0, 0, 0), "f64")
    # translation by V = (0.01, -0.02) in a fluid at rest: the moving wall
    # drags the fluid along V (net momentum of the field has the sign of V)
    V = np.array([0.01, -0.02])
    env.mask = mask
    env.solids = np.array([[20, 15, 1, 0, 20 - V[0], 15 - V[1], 1, 0]], np.float32)
    out = orc.step(env, m, "f64")
    assert (out[0] * out[1]).sum() * V[0] > 0 and (out[0] * out[2]).sum() * V[1] > 0
    # direct check of u_p through the NumPy restatement's wall velocity
    h = np.array([[24.5, 20.0], [15.0, 10.5]])
    up, r = npo._wall_velocity(9, env.solids, np.array([0, 0]), h)
    assert np.allclose(up, V[:, None], atol=1e-6)
    th = 0.01
    env.solids = np.array([[20, 15, np.cos(th), np.sin(th), 20, 15, 1, 0]], np.float64)
    up, r = npo._wall_velocity(9, env.solids, np.array([0, 0]), h)
    q = np.array([[np.cos(th), np.sin(th)], [-np.sin(th), np.cos(th)]]) @ (h - [[20], [15]])
    assert np.allclose(up, (h - [[20], [15]]) - q, atol=1e-12)


def test_symmetric_body_has_no_lift():
    """Disc on the mirror line of a channel with a mirror-symmetric state:
    the y-force and torque vanish (to FP64 rounding) after several steps."""
    nx, ny = 40, 21
    env = orc.Env(9, nx, ny, nu=0.02, bc_type=(0, 1, 1, 1), bc_value=((0.05, 0, 0),) + ((0, 0, 0),) * 5)
    env.mask = orc.voxelize(env, [12, 10, 0, 12, 10, 0, 3.5, 0])
    env.solids = np.array([[12, 10, 1, 0, 12, 10, 1, 0]], np.float32)
    m = orc.init_uniform(env, (0.05, 0, 0), "f64")
    cur = m
    for _ in range(12):
        f = orc.compute_forces(env, cur, "f64")
        cur = orc.step(env, cur, "f64")
    assert f[0, 0] > 1e-5  # drag along the flow
    assert abs(f[0, 1]) < 1e-12 * abs(f[0, 0]) + 1e-15
    assert abs(f[0, 5]) < 1e-12
    assert np.allclose(cur[2], -cur[2][::-1, :], atol=1e-15)  # uy odd in y


def test_mass_fp64():
    env = orc.Env(27, 6, 5, 4)
    m = random_state(27, (4, 5, 6), 3).astype(np.float32)
    assert abs(orc.mass(env, m) - m[0].astype(np.float64).sum()) < 1e-10


def test_polygon_voxelizer_square_closed_form():
    """Axis-aligned square [10.5, 20.5] x [5.5, 12.5]: inside iff
    11 <= x <= 20 and 6 <= y <= 12 (even-odd rule, lbm_func.py:846-859)."""
    env = orc.Env(9, 32, 20)
    sq = np.array([[[10.5, 5.5], [20.5, 5.5], [20.5, 12.5], [10.5, 12.5]]], np.float32)
    m = orc.voxelize_polygons(env, sq, [4])
    y, x = np.mgrid[0:20, 0:32]
    assert np.array_equal(m, ((x >= 11) & (x <= 20) & (y >= 6) & (y <= 12)).astype(np.uint8))


@pytest.mark.parametrize("seed", range(3))
def test_polygon_voxelizer_matches_bruteforce(seed):
    """Random star-shaped polygons: the FP32 even-odd mask equals an FP64
    winding-number test away from the edges; the lowest polygon wins."""
    rng = np.random.default_rng(seed)
    env = orc.Env(9, 48, 40)
    polys, counts = [], []
    for k in range(3):
        n = int(rng.integers(5, 12))
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
        rad = rng.uniform(4, 9, n)
        c = rng.uniform(12, 30, 2)
        polys.append(np.stack([c[0] + rad * np.cos(ang), c[1] + rad * np.sin(ang)], -1))
        counts.append(n)
    verts = np.zeros((3, 12, 2), np.float32)
    for k in range(3):
        verts[k, : counts[k]] = polys[k]
    m = orc.voxelize_polygons(env, verts, counts)
    y, x = np.mgrid[0:40, 0:48].astype(float)
    expect = np.zeros((40, 48), np.uint8)
    ambiguous = np.zeros((40, 48), bool)
    for k in range(2, -1, -1):
        P = verts[k, : counts[k]].astype(float)
        wind = np.zeros((40, 48))
        for j in range(counts[k]):
            a, b = P[j], P[(j + 1) % counts[k]]
            t1 = np.arctan2(a[1] - y, a[0] - x)
            t2 = np.arctan2(b[1] - y, b[0] - x)
            d = (t2 - t1 + np.pi) % (2 * np.pi) - np.pi
            wind += d
            # distance to the edge
            ab = b - a
            t = np.clip(((x - a[0]) * ab[0] + (y - a[1]) * ab[1]) / (ab @ ab), 0, 1)
            dist = np.hypot(x - (a[0] + t * ab[0]), y - (a[1] + t * ab[1]))
            ambiguous |= dist < 1e-3
        inside = np.abs(wind) > np.pi
        expect[inside] = k + 1
    assert np.array_equal(m[~ambiguous], expect[~ambiguous])
    assert (m != 0).sum() > 20

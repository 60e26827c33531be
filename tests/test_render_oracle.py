"""The render oracle (oracle/render.py) pinned by closed forms: linear
velocity fields have exact vorticity (central and one-sided differences agree
for them, so a wrong component, axis, sign or boundary index shows), Pythagorean
speeds, the RGB scaling, and hand-built rays for the projections."""
import numpy as np
import pytest

from oracle import render as R

F = np.float32


def _linear(shape, coef):
    """u_c = sum_a coef[c][a] * x_a (x_0 = x fastest), exact in FP32."""
    nd = len(shape)
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")  # z/y..x order
    coords = grids[::-1]  # x, y, (z)
    u = np.zeros((len(coef),) + tuple(shape))
    for c, row in enumerate(coef):
        for a, k in enumerate(row):
            u[c] += k * coords[a]
    return u.astype(F)


def test_vorticity_2d_linear_field():
    nx, ny = 16, 12
    a, g = 2.0 ** -6, 3 * 2.0 ** -7  # u = a y, v = g x
    u = _linear((ny, nx), [[0, a], [g, 0]])
    img = R.plane(u, "vorticity")
    assert img.shape == (ny, nx)
    assert np.all(img == F((g - a) * nx))


@pytest.mark.parametrize("normal", [2, 0])
def test_vorticity_3d_linear_field(normal):
    nx, ny, nz = 16, 10, 8
    al, be, ga, de, ep, ze = (k * 2.0 ** -8 for k in (1, 3, 5, 7, 9, 11))
    # u = al y + be z, v = ga x + de z, w = ep x + ze y
    u = _linear((nz, ny, nx), [[0, al, be], [ga, 0, de], [ep, ze, 0]])
    img = R.plane(u, "vorticity", normal, 3)
    want = (ga - al) * nx if normal == 2 else (ze - de) * nx
    assert img.shape == ((ny, nx) if normal == 2 else (nz, ny))
    assert np.all(img == F(want))
    # curl about y (front projections): du/dz - dw/dx
    gl = R.grid_length(nx)
    assert np.all(R.curl(u, 1, gl) == F((be - ep) * nx))


def test_speed_and_rgb_closed_forms():
    u = np.zeros((3, 2, 3, 4), F)
    u[0], u[1], u[2] = 2 * 2.0 ** -8, 3 * 2.0 ** -8, 6 * 2.0 ** -8
    assert np.all(R.speed(u) == F(7 * 2.0 ** -8))
    u2 = np.zeros((2, 3, 4), F)
    u2[0], u2[1] = 3 * 2.0 ** -4, -4 * 2.0 ** -4
    rgb = R.velocity_rgb(u2, 0.5)
    assert np.allclose(rgb[..., 0], 0.625) and np.allclose(rgb[..., 1], (0.1875 + 0.5) / 1.0)
    assert np.allclose(rgb[..., 2], (-0.25 + 0.5) / 1.0)
    u2[0] = 4.0  # saturates
    assert np.all(R.velocity_rgb(u2, 0.5)[..., :2] == 1.0)


def test_projections_on_hand_built_rays():
    nz, ny, nx = 7, 6, 5
    u = np.zeros((3, nz, ny, nx), F)
    u[0, 2, 3, 2] = 0.25       # low in z
    u[0, 4, 3, 2] = 0.125      # higher in z: seen first from the top
    u[0, 6, 3, 2] = 0.5        # on the top face: skipped
    top_first = R.projection(u, "first_speed", 2)
    top_max = R.projection(u, "max_speed", 2)
    assert top_first.shape == (ny, nx)
    assert top_first[3, 2] == F(0.125) and top_max[3, 2] == F(0.25)
    assert top_first[0, 2] == 0 and top_max[0, 0] == 0  # face columns
    side = R.projection(u, "max_speed", 0)
    front = R.projection(u, "max_speed", 1)
    assert side.shape == (nz, ny) and side[2, 3] == F(0.25) and side[4, 3] == F(0.125)
    assert front.shape == (nz, nx) and front[2, 2] == F(0.25)
    # a ray through the mask reads 1000
    mask = np.zeros((nz, ny, nx), np.uint8)
    mask[1, 2, 1] = 1
    assert R.projection(u, "max_speed", 2, mask)[2, 1] == 1000
    assert R.projection(u, "max_speed", 0, mask)[1, 2] == 1000


def test_vorticity_projection_keeps_sign_of_largest():
    nz, ny, nx = 8, 8, 8
    u = np.zeros((3, nz, ny, nx), F)
    gl = R.grid_length(nx)
    u[1, 2, 3, 4], u[1, 2, 3, 2] = 1.0, -1.0     # dv/dx at (3, 3, 2) = 2 / (2 gl)
    u[1, 5, 3, 2], u[1, 5, 3, 4] = 2.0, -2.0     # at z = 5: -4 / (2 gl): larger, negative
    img = R.projection(u, "max_vorticity", 2)
    assert img[3, 3] == F(-4.0) / (F(2.0) * gl)


def test_solid_near_matches_definition():
    rng = np.random.default_rng(5)
    m = (rng.random((9, 11)) < 0.05).astype(np.uint8)
    for t in (0.0, 1.0, 1.5, 2.0):
        got = R.solid_near(m, t)
        ys, xs = np.nonzero(m)
        want = m != 0
        for y in range(9):
            for x in range(11):
                if t > 0 and len(ys):
                    d2 = (ys - y) ** 2 + (xs - x) ** 2
                    want[y, x] |= bool((d2 <= t * t).any())
        assert np.array_equal(got, want), t

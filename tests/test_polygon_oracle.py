"""The exact polygon cut-cell step of the oracle (hlo_step_polygons, the
reference's get_cutcell, envs/lbm/lbm_func.py:545-607, and its cut branch,
lbm_func.py:346-401) pinned independently of its own formulas:

* the set of cut links, the nearest crossing and the crossing point come
  from an exact rational segment-intersection test; on a rest state every
  cut link i pushes -2 w_i c_i and a torque (P - position) x F / nx, so the
  per-solid forces are closed forms of that exact classification;
* a static rectangle whose link cut set equals its mask's gives the mask
  rule's oracle step bit for bit (the wall rule with u_p = 0);
* a translating rectangle gives the mask rule's step within rounding
  (the material-point velocity equals the translation);
* with no cell inside a bounding circle the step is the plain step.
"""
from fractions import Fraction as Fr

import numpy as np

from oracle import oracle as orc
from tests.helpers import random_state

CX = (0, 1, -1, 0, 0, 1, -1, -1, 1)
CY = (0, 0, 0, 1, -1, 1, 1, -1, -1)
W32 = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36], np.float32)


def _env(nx, ny, pos, nu=0.02):
    env = orc.Env(9, nx, ny, nu=nu, bc_type=(0, 1, 1, 1), bc_value=((0.03, 0, 0),) + ((0, 0, 0),) * 5)
    P = np.zeros((len(pos), 8), np.float32)
    for k, (px, py) in enumerate(pos):
        P[k] = (px, py, 1, 0, px, py, 1, 0)
    env.solids = P
    return env


def _tables(polys, maxv=None):
    maxv = maxv or max(len(p) for p in polys)
    v = np.zeros((len(polys), maxv, 2), np.float32)
    for k, p in enumerate(polys):
        v[k, : len(p)] = p
    return v, [len(p) for p in polys]


def _orient(p, q, r):
    return (p[0] - r[0]) * (q[1] - r[1]) - (p[1] - r[1]) * (q[0] - r[0])


def exact_cuts(nx, ny, polys32, pos32, r2_32):
    """{(x, y, i): (solid, crossing point)} by exact rational arithmetic:
    link a = b - c_i -> b crosses segment c -> d when c and d lie strictly on
    opposite sides of line ab and a, b not strictly on one side of line cd;
    the crossing nearest b wins, the first one on a tie."""
    F = lambda v: Fr(float(v))  # noqa: E731
    polys = [[(F(x), F(y)) for x, y in p] for p in polys32]
    pos = [(F(x), F(y)) for x, y in pos32]
    r2 = [F(r) for r in r2_32]
    out = {}
    for y in range(1, ny - 1):
        for x in range(1, nx - 1):
            b = (Fr(x), Fr(y))
            near = [s for s in range(len(polys)) if (b[0] - pos[s][0]) ** 2 + (b[1] - pos[s][1]) ** 2 <= r2[s]]
            if not near:
                continue
            for i in range(1, 9):
                a = (b[0] - CX[i], b[1] - CY[i])
                best = None
                for s in near:
                    P = polys[s]
                    for n in range(len(P)):
                        c, d = P[n], P[(n + 1) % len(P)]
                        abc, abd = _orient(a, b, c), _orient(a, b, d)
                        if abc * abd >= 0:
                            continue
                        cda, cdb = _orient(c, d, a), _orient(c, d, b)
                        if cda * cdb > 0:
                            continue
                        tau = cda / (cda - cdb)  # position along a -> b
                        sr = abc / (abc - abd)  # position along c -> d
                        if best is None or tau > best[0]:
                            best = (tau, s, (c[0] + (d[0] - c[0]) * sr, c[1] + (d[1] - c[1]) * sr))
                if best is not None:
                    out[(x, y, i)] = (best[1], best[2])
    return out


def rest_forces(cuts, nsolid, pos32, nx):
    F = np.zeros((nsolid, 3))
    for (x, y, i), (s, p) in cuts.items():
        fx = -2.0 * float(W32[i]) * CX[i]
        fy = -2.0 * float(W32[i]) * CY[i]
        rx, ry = float(p[0]) - float(pos32[s][0]), float(p[1]) - float(pos32[s][1])
        F[s] += (fx, fy, (rx * fy - ry * fx) / nx)
    return F


def _rest(nx, ny):
    m = np.zeros((6, ny, nx), np.float32)
    m[0] = 1.0
    return m


def test_rest_state_forces_match_exact_cut_links():
    nx, ny = 40, 34
    tri = np.array([[11.3, 9.7], [24.9, 13.2], [14.1, 23.6]], np.float32)
    hexa = np.array([[29.2, 20.1], [33.7, 22.4], [33.9, 27.3], [29.6, 29.8], [25.3, 27.1], [25.7, 22.2]], np.float32)
    pos = [(16.0, 15.0), (29.5, 25.0)]
    env = _env(nx, ny, pos)
    v, cnt = _tables([tri, hexa])
    r2 = np.array([12.0 ** 2, 7.0 ** 2], np.float32)
    m = _rest(nx, ny)
    out, F = orc.step_polygons(env, m, v, v, cnt, r2)
    cuts = exact_cuts(nx, ny, [tri, hexa], pos, r2)
    want = rest_forces(cuts, 2, pos, nx)
    assert len(cuts) > 100 and {s for s, _ in cuts.values()} == {0, 1}
    np.testing.assert_array_equal(F[:, 0:2], want[:, 0:2])  # sums of -2 w c are exact
    np.testing.assert_allclose(F[:, 5], want[:, 2], rtol=1e-5, atol=1e-7)
    assert np.array_equal(out, orc.step(env, m))  # at rest a wall population equals the pulled one


def test_nearest_crossing_and_bounding_circle():
    """Two thin walls between x = 12 and x = 13: a link crossing both is
    charged to the wall nearer the cell; a small bounding circle limits the
    cut cells."""
    nx, ny = 30, 26
    w0 = np.array([[12.3, 5.2], [12.5, 5.2], [12.5, 20.9], [12.3, 20.9]], np.float32)
    w1 = np.array([[12.7, 5.6], [12.9, 5.6], [12.9, 21.3], [12.7, 21.3]], np.float32)
    pos = [(12.4, 13.0), (12.8, 13.5)]
    env = _env(nx, ny, pos)
    v, cnt = _tables([w0, w1])
    for r2 in (np.array([100.0, 100.0], np.float32), np.array([9.0, 20.0], np.float32)):
        _, F = orc.step_polygons(env, _rest(nx, ny), v, v, cnt, r2)
        cuts = exact_cuts(nx, ny, [w0, w1], pos, r2)
        want = rest_forces(cuts, 2, pos, nx)
        np.testing.assert_array_equal(F[:, 0:2], want[:, 0:2])
        np.testing.assert_allclose(F[:, 5], want[:, 2], rtol=1e-5, atol=1e-7)
    # the links between x = 12 and x = 13 cross both walls (four edges)
    big = exact_cuts(nx, ny, [w0, w1], pos, np.array([100.0, 100.0], np.float32))
    assert big[(13, 10, 1)][0] == 1 and big[(12, 10, 2)][0] == 0
    assert float(big[(13, 10, 1)][1][0]) == float(np.float32(12.9))


def _rect(x0, y0, x1, y1):
    return np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]], np.float32)


def test_static_rectangle_equals_mask_rule():
    """Edge offsets (left .625, right .25, bottom .875, top .5) such that no
    diagonal link cuts or meets a corner: every link crosses the boundary at most
    once, so the cut links are the mask's, and with u_p = 0 both rules give
    the same numbers."""
    nx, ny = 36, 30
    rect = _rect(10.625, 8.875, 21.25, 17.5)
    pos = [(15.875, 13.25)]
    env = _env(nx, ny, pos)
    v, cnt = _tables([rect])
    r2 = np.array([200.0], np.float32)
    m = random_state(9, (ny, nx), seed=3).astype(np.float32)
    got, F = orc.step_polygons(env, m, v, v, cnt, r2)
    env.mask = orc.voxelize_polygons(env, v, cnt)
    want = orc.step(env, m)
    Fm = orc.compute_forces(env, m)
    cuts = exact_cuts(nx, ny, [rect], pos, r2)
    mask_links = {(x, y, i) for y in range(1, ny - 1) for x in range(1, nx - 1) for i in range(1, 9)
                  if (env.mask[y, x] != 0) != (env.mask[y - CY[i], x - CX[i]] != 0)}
    assert set(cuts) == mask_links and len(mask_links) > 50
    assert np.array_equal(got, want)
    np.testing.assert_array_equal(F[:, 0:2], Fm[:, 0:2])


def test_translating_rectangle_matches_mask_rule_within_rounding():
    nx, ny = 36, 30
    d = np.float32(0.25)
    rect = _rect(10.625, 8.875, 21.25, 17.5)
    last = rect - np.array([d, 0], np.float32)
    pos = [(15.875, 13.25)]
    env = _env(nx, ny, pos)
    env.solids[0, 4] = np.float32(15.875) - d  # last position: u_p = (0.25, 0) everywhere
    v, cnt = _tables([rect])
    vl, _ = _tables([last])
    m = random_state(9, (ny, nx), seed=4).astype(np.float32)
    got, F = orc.step_polygons(env, m, v, vl, cnt, np.array([200.0], np.float32))
    env.mask = orc.voxelize_polygons(env, v, cnt)
    want = orc.step(env, m)
    Fm = orc.compute_forces(env, m)
    moved = np.abs(got - orc.step(_env(nx, ny, pos), m)).max()
    assert moved > 1e-3  # the body matters
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-6)
    np.testing.assert_allclose(F[:, 0:2], Fm[:, 0:2], rtol=1e-5)


def test_no_cell_in_a_circle_is_the_plain_step():
    nx, ny = 24, 20
    sq = _rect(8.3, 7.6, 12.3, 11.6)
    env = _env(nx, ny, [(10.3, 9.6)])
    v, cnt = _tables([sq])
    m = random_state(9, (ny, nx), seed=5).astype(np.float32)
    got, F = orc.step_polygons(env, m, v, v, cnt, np.array([0.01], np.float32))  # no lattice point within 0.1
    assert np.array_equal(got, orc.step(env, m)) and not F.any()

"""The planar chain oracle (oracle/chain.py) and the MJCF chain reader.

The oracle restates MuJoCo's forward dynamics for planar trees (the engine
the reference's coupled loop steps, envs/lbm/lbm_fluid_env.py:441-533). It is
pinned here by what the mechanics fixes, independently of its own formulas:
* the textbook two-link inertia matrix, Coriolis and gravity terms;
* the kinetic energy of bodies made of point masses, from finite-difference
  velocities of the points (pins M, Jacobians and inertia aggregation);
* the Christoffel form of the bias force from finite differences of M;
* linear and angular momentum rates equal to the external loads;
* a damped one-dof closed form (RK4) and the semi-implicit Euler recurrence;
* constant-force motion, exact under RK4, with the reference's doubled
  generalized load (xfrc_applied plus its accumulation in qfrc_applied).
The reader is pinned by closed-form capsule volumes, centres of mass, an
independent quadrature of the capsule inertia, and eel_2d.xml's attributes.
"""
import math
import os
import textwrap

import numpy as np
import pytest
from scipy.integrate import quad

from oracle import chain as oc


def to_oracle(m) -> oc.ChainModel:
    return oc.ChainModel(
        parent=m.parent, free=m.free, pos=m.pos, angle=m.angle, com=m.com, mass=m.mass, inertia=m.inertia,
        posz=m.posz, comz=m.comz, dof_body=m.dof_body, dof_type=m.dof_type, dof_axis=m.dof_axis,
        dof_anchor=m.dof_anchor, dof_sign=m.dof_sign, damping=m.damping, stiffness=m.stiffness,
        springref=m.springref, armature=m.armature, qpos0=m.qpos0, act_dof=m.act_dof, gear=m.gear,
        ctrl_lo=m.ctrl_lo, ctrl_hi=m.ctrl_hi, ctrl_limited=m.ctrl_limited, gravity=np.asarray(m.gravity),
        dt=m.dt, integrator=m.integrator)


def model(nb, dofs, parent, pos=None, angle=None, com=None, mass=None, inertia=None, free=None, **kw):
    """dofs: list of (body, type, axis, anchor, sign)."""
    nv = len(dofs)
    z = np.zeros
    m = oc.ChainModel(
        parent=np.array(parent), free=np.array(free if free is not None else [False] * nb),
        pos=np.array(pos if pos is not None else z((nb, 2)), float).reshape(nb, 2),
        angle=np.array(angle if angle is not None else z(nb), float),
        com=np.array(com if com is not None else z((nb, 2)), float).reshape(nb, 2),
        mass=np.array(mass if mass is not None else np.ones(nb), float),
        inertia=np.array(inertia if inertia is not None else np.ones(nb) * 0.1, float),
        posz=z(nb), comz=z(nb),
        dof_body=np.array([d[0] for d in dofs]), dof_type=np.array([d[1] for d in dofs]),
        dof_axis=np.array([d[2] for d in dofs], float).reshape(nv, 2),
        dof_anchor=np.array([d[3] for d in dofs], float).reshape(nv, 2),
        dof_sign=np.array([d[4] for d in dofs], float),
        damping=np.array(kw.pop("damping", z(nv)), float), stiffness=np.array(kw.pop("stiffness", z(nv)), float),
        springref=z(nv), armature=np.array(kw.pop("armature", z(nv)), float), qpos0=z(nv),
        act_dof=np.array(kw.pop("act_dof", []), int), gear=np.array(kw.pop("gear", []), float),
        ctrl_lo=np.array(kw.pop("ctrl_lo", []), float), ctrl_hi=np.array(kw.pop("ctrl_hi", []), float),
        ctrl_limited=np.array(kw.pop("ctrl_limited", []), bool), gravity=np.array(kw.pop("gravity", (0.0, 0.0))),
        dt=kw.pop("dt", 0.01), integrator=kw.pop("integrator", "RK4"))
    assert not kw
    return m


H, S = oc.HINGE, oc.SLIDE


def two_link(l1=0.7, lc1=0.3, lc2=0.45, m1=1.3, m2=0.8, I1=0.05, I2=0.03, **kw):
    """Fixed base: hinge at the origin, link 1 along -y; hinge 2 at (0, -l1)."""
    return model(2, [(0, H, (0, 0), (0, 0), 1.0), (1, H, (0, 0), (0, 0), 1.0)], [-1, 0],
                 pos=[(0, 0), (0, -l1)], com=[(0, -lc1), (0, -lc2)], mass=[m1, m2], inertia=[I1, I2], **kw)


def tree():
    """Free planar root (x, y, rz), a slide child, a hinge child with an
    offset anchor, and a grandchild on the hinge child: the dof types,
    anchors, body angles and a branch together."""
    dofs = [(0, S, (1, 0), (0, 0), 1.0), (0, S, (0, 1), (0, 0), 1.0), (0, H, (0, 0), (0, 0), 1.0),
            (1, S, (0.6, 0.8), (0, 0), 1.0),
            (2, H, (0, 0), (0.1, -0.05), -1.0),
            (3, H, (0, 0), (0.02, 0.03), 1.0)]
    return model(4, dofs, [-1, 0, 0, 2], free=[True, False, False, False],
                 pos=[(0, 0), (0.3, 0.1), (-0.2, -0.4), (0.05, -0.5)], angle=[0, 0.4, -0.3, 0.7],
                 com=[(0.05, -0.02), (0.1, 0.0), (0.03, -0.2), (-0.04, -0.1)],
                 mass=[1.1, 0.4, 0.7, 0.5], inertia=[0.02, 0.004, 0.03, 0.01])


# ------------------------------------------------------------- kinematics

def test_two_link_kinematics_closed_form():
    m = two_link()
    q = np.array([0.4, -1.1])
    K = oc.kinematics(m, q, np.zeros(2))
    d1 = np.array([math.sin(q[0]), -math.cos(q[0])])
    d2 = np.array([math.sin(q[0] + q[1]), -math.cos(q[0] + q[1])])
    np.testing.assert_allclose(K.c[0], 0.3 * d1, atol=1e-15)
    np.testing.assert_allclose(K.p[1], 0.7 * d1, atol=1e-15)
    np.testing.assert_allclose(K.c[1], 0.7 * d1 + 0.45 * d2, atol=1e-15)


def test_anchor_and_slide_kinematics():
    # one body: slide along (0.6, 0.8) then a hinge about anchor a = (0.2, 0.1)
    m = model(1, [(0, S, (0.6, 0.8), (0, 0), 1.0), (0, H, (0, 0), (0.2, 0.1), 1.0)], [-1],
              pos=[(1.0, 2.0)], angle=[0.5], com=[(0.3, -0.1)])
    q = np.array([0.7, -0.9])
    K = oc.kinematics(m, q, np.zeros(2))
    R = lambda a: np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])
    o = np.array([1.0, 2.0]) + R(0.5) @ np.array([0.6, 0.8]) * 0.7  # origin after the slide
    A = o + R(0.5) @ np.array([0.2, 0.1])                            # the anchor stays fixed
    p = A - R(0.5 - 0.9) @ np.array([0.2, 0.1])
    np.testing.assert_allclose(K.p[0], p, atol=1e-15)
    np.testing.assert_allclose(K.c[0], p + R(-0.4) @ np.array([0.3, -0.1]), atol=1e-15)


# ------------------------------------------------------------- dynamics

def test_two_link_matches_textbook():
    """M, Coriolis and gravity of the textbook planar two-link arm (relative
    joint angles): M11 = m1 lc1^2 + m2 (l1^2 + lc2^2 + 2 l1 lc2 cos q2) + I1 + I2, ..."""
    l1, lc1, lc2, m1, m2, I1, I2, g = 0.7, 0.3, 0.45, 1.3, 0.8, 0.05, 0.03, 9.81
    m = two_link(gravity=(0.0, -g))
    rng = np.random.default_rng(3)
    for _ in range(5):
        q, qd = rng.uniform(-3, 3, 2), rng.uniform(-2, 2, 2)
        c2, s2 = math.cos(q[1]), math.sin(q[1])
        M = np.array([[m1 * lc1 ** 2 + m2 * (l1 ** 2 + lc2 ** 2 + 2 * l1 * lc2 * c2) + I1 + I2,
                       m2 * (lc2 ** 2 + l1 * lc2 * c2) + I2],
                      [m2 * (lc2 ** 2 + l1 * lc2 * c2) + I2, m2 * lc2 ** 2 + I2]])
        np.testing.assert_allclose(oc.mass_matrix(m, q), M, rtol=1e-13)
        h = -m2 * l1 * lc2 * s2
        cor = np.array([h * qd[1] * (2 * qd[0] + qd[1]), -h * qd[0] ** 2])
        # links hang along -y: V = -m1 g lc1 cos q1 - m2 g (l1 cos q1 + lc2 cos(q1 + q2))
        G = np.array([-m1 * g * lc1 * math.sin(q[0]) - m2 * g * (l1 * math.sin(q[0]) + lc2 * math.sin(q[0] + q[1])),
                      -m2 * g * lc2 * math.sin(q[0] + q[1])])
        acc = oc.forward(m, q, qd, [], np.zeros((2, 6)), np.zeros(2))[0]
        np.testing.assert_allclose(acc, np.linalg.solve(M, G - cor), rtol=1e-11, atol=1e-12)


def _point_bodies(rng, m, npts=7):
    """Replace each body's mass properties by those of random point masses."""
    pts = []
    for b in range(m.nbody):
        r = rng.uniform(-0.3, 0.3, (npts, 2))
        w = rng.uniform(0.1, 1.0, npts)
        c = (w[:, None] * r).sum(0) / w.sum()
        m.mass[b] = w.sum()
        m.com[b] = c
        m.inertia[b] = (w * ((r - c) ** 2).sum(1)).sum()
        pts.append((r, w))
    return pts


def _world_points(m, q, pts):
    K = oc.kinematics(m, q, np.zeros(m.nv))
    out = []
    for b, (r, w) in enumerate(pts):
        c, s = K.cs[b]
        out.append(K.p[b] + r @ np.array([[c, s], [-s, c]]))
    return out


def test_kinetic_energy_of_point_masses():
    rng = np.random.default_rng(11)
    m = tree()
    pts = _point_bodies(rng, m)
    for _ in range(4):
        q, qd = rng.uniform(-1.5, 1.5, m.nv), rng.uniform(-2, 2, m.nv)
        eps = 1e-6
        xp, xm = _world_points(m, q + eps * qd, pts), _world_points(m, q - eps * qd, pts)
        T = sum(0.5 * (w * (((a - b) / (2 * eps)) ** 2).sum(1)).sum() for (_, w), a, b in zip(pts, xp, xm))
        np.testing.assert_allclose(oc.energy(m, q, qd), T, rtol=1e-8)


def test_bias_is_christoffel_form_of_mass_matrix():
    rng = np.random.default_rng(5)
    m = tree()
    n = m.nv
    for _ in range(3):
        q, qd = rng.uniform(-1.5, 1.5, n), rng.uniform(-2, 2, n)
        e = 1e-6
        dM = np.stack([(oc.mass_matrix(m, q + e * np.eye(n)[k]) - oc.mass_matrix(m, q - e * np.eye(n)[k])) / (2 * e)
                       for k in range(n)])  # dM[k] = dM/dq_k
        c = np.array([sum((dM[k][i, j] - 0.5 * dM[i][j, k]) * qd[j] * qd[k] for j in range(n) for k in range(n))
                      for i in range(n)])
        bias = oc.forward(m, q, qd, [], np.zeros((m.nbody, 6)), np.zeros(n))[2]
        np.testing.assert_allclose(bias, c, rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("factor", [0.0, 1.0])
def test_momentum_rates_equal_the_loads(factor):
    rng = np.random.default_rng(8)
    m = tree()
    xfrc = np.zeros((m.nbody, 6))
    xfrc[:, 0:2] = rng.normal(size=(m.nbody, 2))
    xfrc[:, 5] = rng.normal(size=m.nbody)
    for _ in range(3):
        q, qd = rng.uniform(-1.5, 1.5, m.nv), rng.uniform(-2, 2, m.nv)
        K = oc.kinematics(m, q, qd)
        J, Jp = oc.jacobians(m, K)
        applied = factor * oc.xfrc_generalized(m, J, Jp, xfrc)
        acc, _, _, J, Jp = oc.forward(m, q, qd, [], xfrc, applied)
        a = oc.bias_accel(m, K, qd)
        cdd = np.array([J[b] @ acc + a[b] for b in range(m.nbody)])
        phidd = Jp @ acc
        scale = 1.0 + factor
        np.testing.assert_allclose((m.mass[:, None] * cdd).sum(0), scale * xfrc[:, 0:2].sum(0), rtol=1e-11,
                                   atol=1e-12)
        cross = lambda u, v: u[0] * v[1] - u[1] * v[0]
        Ldot = sum(m.mass[b] * cross(K.c[b], cdd[b]) + m.inertia[b] * phidd[b] for b in range(m.nbody))
        tau = sum(cross(K.c[b], xfrc[b, 0:2]) + xfrc[b, 5] for b in range(m.nbody))
        assert Ldot == pytest.approx(scale * tau, rel=1e-10, abs=1e-12)


def test_damped_hinge_closed_form_rk4():
    """(I + m lc^2 + armature) th'' + d th' = gear * ctrl from rest."""
    m = model(1, [(0, H, (0, 0), (0, 0), 1.0)], [-1], com=[(0, -0.2)], mass=[0.9], inertia=[0.02],
              damping=[0.04], armature=[0.01], act_dof=[0], gear=[0.12], ctrl_lo=[-1], ctrl_hi=[1],
              ctrl_limited=[True])
    J, d, tau = 0.02 + 0.9 * 0.04 + 0.01, 0.04, 0.12
    q, v = np.zeros(1), np.zeros(1)
    for _ in range(150):
        q, v, _ = oc.step(m, q, v, [3.0], np.zeros((1, 6)))  # ctrl clamped to 1
    t = 150 * 0.01
    assert v[0] == pytest.approx(tau / d * (1 - math.exp(-d * t / J)), rel=1e-9)
    assert q[0] == pytest.approx(tau / d * (t - J / d * (1 - math.exp(-d * t / J))), rel=1e-9)


def test_euler_recurrence_with_implicit_damping():
    m = model(1, [(0, H, (0, 0), (0, 0), 1.0)], [-1], com=[(0, -0.2)], mass=[0.9], inertia=[0.02],
              damping=[0.04], armature=[0.01], act_dof=[0], gear=[0.12], ctrl_lo=[-1], ctrl_hi=[1],
              ctrl_limited=[True], integrator="Euler")
    J, d, tau, h = 0.02 + 0.9 * 0.04 + 0.01, 0.04, -0.06, 0.01
    q, v = np.zeros(1), np.zeros(1)
    qe, ve = 0.0, 0.0
    for _ in range(40):
        q, v, _ = oc.step(m, q, v, [-0.5], np.zeros((1, 6)))
        ve = (J * ve + h * tau) / (J + h * d)  # (J + h d) a = tau - d v; v += h a
        qe = qe + h * ve
    assert v[0] == pytest.approx(ve, rel=1e-12) and q[0] == pytest.approx(qe, rel=1e-12)


@pytest.mark.parametrize("factor", [0.0, 1.0])
def test_constant_force_motion_is_exact(factor):
    """A free body (centre of mass at its origin) under a constant force and
    torque: x = F t^2 / (2 (m + armature)), th = T t^2 / (2 (I + armature)),
    both doubled when the load also enters qfrc_applied."""
    m = model(1, [(0, S, (1, 0), (0, 0), 1.0), (0, S, (0, 1), (0, 0), 1.0), (0, H, (0, 0), (0, 0), 1.0)], [-1],
              free=[True], mass=[2.0], inertia=[0.3], armature=[0.01, 0.01, 0.01])
    xfrc = np.zeros((1, 6))
    xfrc[0, 0], xfrc[0, 1], xfrc[0, 5] = 0.5, -0.25, 0.2
    q, v = np.zeros(3), np.zeros(3)
    for _ in range(20):
        q, v, applied = oc.step(m, q, v, [], xfrc, applied_factor=factor)
    t, s = 0.2, 1.0 + factor
    np.testing.assert_allclose(q, s * np.array([0.5 / 2.01, -0.25 / 2.01, 0.2 / 0.31]) * t * t / 2, rtol=1e-12)
    np.testing.assert_allclose(applied, factor * np.array([0.5, -0.25, 0.2]), rtol=1e-15)


def test_actuators_clamp_and_gear():
    m = model(1, [(0, H, (0, 0), (0, 0), 1.0)], [-1], act_dof=[0, 0], gear=[2.0, 0.5], ctrl_lo=[-1, 0],
              ctrl_hi=[1, 4], ctrl_limited=[True, False])
    np.testing.assert_allclose(oc.actuator_force(m, [5.0, -3.0]), [2.0 * 1 + 0.5 * -3.0])


# ------------------------------------------------------------- MJCF reader

EEL = os.path.join(os.path.dirname(__file__), "golden", "eel_2d.xml")


def _capsule_izz_quadrature(r, h, rho):
    """Moment of a capsule (axis length h, radius r) about an axis through
    its centre perpendicular to its axis, by slicing into disks."""
    rad = lambda s: r if abs(s) <= h / 2 else math.sqrt(max(r * r - (abs(s) - h / 2) ** 2, 0.0))
    f = lambda s: rho * (math.pi * rad(s) ** 4 / 4 + math.pi * rad(s) ** 2 * s * s)
    return sum(quad(f, a, b, epsabs=0, epsrel=1e-13)[0] for a, b in ((-h / 2 - r, -h / 2), (-h / 2, h / 2),
                                                                    (h / 2, h / 2 + r)))


def test_capsule_inertia_formula_against_quadrature():
    from openhomelbm_amd.chain import shape_inertia, shape_volume
    for r, h in ((0.035, 0.09), (0.02, 0.1), (0.5, 0.0), (0.1, 2.0)):
        vol = shape_volume("capsule", np.array([r, h / 2, 0]))
        assert vol == pytest.approx(quad(lambda s: math.pi * (r if abs(s) <= h / 2 else
                                                              math.sqrt(max(r * r - (abs(s) - h / 2) ** 2, 0))) ** 2,
                                         -h / 2 - r, h / 2 + r, points=[-h / 2, h / 2])[0], rel=1e-10)
        I = shape_inertia("capsule", np.array([r, h / 2, 0]), 450.0 * vol)
        assert I[0] == pytest.approx(_capsule_izz_quadrature(r, h, 450.0), rel=1e-10)


def test_eel_model_from_mjcf():
    from openhomelbm_amd.chain import chain_from_mjcf
    m = chain_from_mjcf(EEL)
    assert m.body_names == ["seg1", "seg2", "seg3", "seg4", "seg5", "seg6"]
    assert m.nbody == 6 and m.nv == 8 and m.nact == 5
    assert list(m.parent) == [-1, 0, 1, 2, 3, 4] and list(m.free) == [True] + [False] * 5
    assert list(m.dof_type) == [oc.SLIDE, oc.SLIDE, oc.HINGE] + [oc.HINGE] * 5
    assert m.joint_names[3:] == ["joint1", "joint2", "joint3", "joint4", "joint5"]
    np.testing.assert_array_equal(m.damping, 0.04)
    np.testing.assert_array_equal(m.armature, 0.01)
    np.testing.assert_array_equal(m.gear, [0.12, 0.14, 0.16, 0.18, 0.20])
    assert list(m.act_dof) == [3, 4, 5, 6, 7] and m.ctrl_limited.all()
    np.testing.assert_array_equal(m.ctrl_lo, -1.0)
    assert m.dt == 0.01 and m.integrator == "RK4"
    np.testing.assert_array_equal(m.gravity, [0.0, 0.0])
    np.testing.assert_allclose(m.dof_range[3:, 1], np.radians([50, 55, 60, 65, 70]))
    radii, lengths = [0.035, 0.034, 0.032, 0.029, 0.025, 0.020], [0.09] * 5 + [0.10]
    for b, (r, L) in enumerate(zip(radii, lengths)):
        vol = math.pi * r * r * L + 4.0 / 3.0 * math.pi * r ** 3
        assert m.mass[b] == pytest.approx(450.0 * vol, rel=1e-12)
        np.testing.assert_allclose(m.com[b], [0.0, -L / 2], atol=1e-15)
        assert m.inertia[b] == pytest.approx(_capsule_izz_quadrature(r, L, 450.0), rel=1e-10)
    np.testing.assert_allclose(m.pos[1:], [[0.0, -0.09]] * 5, atol=1e-15)
    np.testing.assert_allclose(m.posz, 0.05)
    np.testing.assert_array_equal(m.qpos0, 0.0)


def test_mjcf_defaults_classes_and_units(tmp_path):
    from openhomelbm_amd.chain import chain_from_mjcf
    xml = textwrap.dedent("""\
        <mujoco>
          <compiler angle="radian"/>
          <option timestep="0.005" gravity="0 -9.81 0"/>
          <default>
            <joint damping="0.5" armature="0.1"/>
            <geom density="100"/>
            <default class="stiff">
              <joint stiffness="2" damping="0.25"/>
            </default>
          </default>
          <worldbody>
            <body name="a" pos="1 2 0" euler="0 0 0.3">
              <joint name="j0" type="hinge" axis="0 0 -1" pos="0.1 0 0" ref="0.2"/>
              <geom type="box" size="0.1 0.2 0.05"/>
              <body name="b" pos="0 -0.5 0" childclass="stiff">
                <joint name="j1" type="slide" axis="1 0 0"/>
                <geom type="sphere" size="0.1" mass="3"/>
              </body>
            </body>
          </worldbody>
          <actuator><motor joint="j1" gear="5"/></actuator>
        </mujoco>""")
    p = tmp_path / "m.xml"
    p.write_text(xml)
    m = chain_from_mjcf(str(p))
    assert m.integrator == "EULER" and m.dt == 0.005
    np.testing.assert_allclose(m.gravity, [0.0, -9.81])
    assert m.angle[0] == pytest.approx(0.3) and list(m.dof_sign) == [-1.0, 1.0]
    np.testing.assert_allclose(m.dof_anchor[0], [0.1, 0.0])
    assert m.qpos0[0] == pytest.approx(0.2)
    np.testing.assert_allclose(m.damping, [0.5, 0.25])
    np.testing.assert_allclose(m.stiffness, [0.0, 2.0])
    np.testing.assert_allclose(m.armature, [0.1, 0.1])
    assert m.mass[0] == pytest.approx(100 * 8 * 0.1 * 0.2 * 0.05) and m.mass[1] == 3.0
    assert m.inertia[0] == pytest.approx(m.mass[0] / 3 * (0.1 ** 2 + 0.2 ** 2))
    assert m.inertia[1] == pytest.approx(0.4 * 3 * 0.01)
    assert not m.ctrl_limited[0] and m.gear[0] == 5.0

"""CPU restatement of the planar articulated-chain step (test infrastructure only).

The reference's coupled environments hand the fluid loads to MuJoCo-Warp and
take one engine step per lattice step (envs/lbm/lbm_fluid_env.py:441-533;
envs/lbm3d/lbm_fluid_env_3d.py:679-770), on MJCF chains such as
envs/lbm/eel/eel_2d.xml: a free root and hinges about z, motors with gears,
joint damping and armature from the defaults, option integrator RK4. MuJoCo
(the `mujoco` / `mujoco-warp` requirements, unpinned in requirements.txt) is
not importable here, so this module restates the published algorithm for a
planar kinematic tree, in FP64:

* forward kinematics of the tree (body offset, then the body's joints in
  order: slides along an axis of the frame before them, hinges about an
  anchor point), with each frame origin's velocity;
* the centre-of-mass Jacobians J_b and angle Jacobians; the joint-space
  inertia M = sum_b m_b J_b^T J_b + I_b Jphi_b^T Jphi_b + diag(armature) (the
  matrix MuJoCo's composite-rigid-body pass builds);
* the bias force sum_b m_b J_b^T (a_b - g), with a_b the centre-of-mass
  acceleration at zero joint acceleration (MuJoCo's recursive Newton-Euler at
  qacc = 0, gravity included);
* passive forces -damping qvel - stiffness (qpos - springref), motor forces
  gear * clamp(ctrl), and the spatial loads xfrc_applied (force at the body's
  centre of mass, torque about z) mapped through J^T at every stage, plus
  qfrc_applied: the same mapping at the step's start configuration times
  `applied_factor` (the reference's loop zeroes qfrc_applied and accumulates
  xfrc_applied into it before each engine step, lbm_fluid_env.py:505-521,
  and the engine adds both);
* MuJoCo's classic four-stage Runge-Kutta (weights 1/6, 1/3, 1/3, 1/6; loads
  and controls held over the step) or its semi-implicit Euler with implicit
  joint damping ((M + dt D) qacc = f; qvel += dt qacc; qpos += dt qvel).

A free joint in the plane is three dofs of its body: slides along world x and
y and a hinge about the body origin; its qpos0 is the body's position and
angle (MuJoCo's free-joint qpos is absolute, so the body offset is not
applied again). The planar reduction is exact when every hinge axis is z,
the loads are planar and each body's inertia has no x-z or y-z products.

Pinned in tests/test_chain_oracle.py by the textbook two-link inertia and
Coriolis terms, the kinetic energy of sampled point masses, the Christoffel
form of the bias from finite differences of M, momentum and energy
invariants, a damped one-dof closed form and constant-force motion.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

SLIDE, HINGE = 0, 1


@dataclass
class ChainModel:
    """A planar kinematic tree. Bodies in tree order (parent < child); each
    body's dofs consecutive, in joint order."""

    parent: np.ndarray        # [nb] int, -1 = world
    free: np.ndarray          # [nb] bool: the body's offset is absorbed in its dofs' qpos0
    pos: np.ndarray           # [nb, 2] body offset in the parent frame
    angle: np.ndarray         # [nb] body rotation about z relative to the parent
    com: np.ndarray           # [nb, 2] centre of mass in the body frame
    mass: np.ndarray          # [nb]
    inertia: np.ndarray       # [nb] zz moment about the centre of mass
    posz: np.ndarray          # [nb] world z of the body origin (constant in the plane)
    comz: np.ndarray          # [nb] world z of the centre of mass
    dof_body: np.ndarray      # [nv] int
    dof_type: np.ndarray      # [nv] SLIDE / HINGE
    dof_axis: np.ndarray      # [nv, 2] slide axis in the frame before the joint
    dof_anchor: np.ndarray    # [nv, 2] hinge anchor in the body frame
    dof_sign: np.ndarray      # [nv] hinge axis +z (1) or -z (-1)
    damping: np.ndarray       # [nv]
    stiffness: np.ndarray     # [nv]
    springref: np.ndarray     # [nv]
    armature: np.ndarray      # [nv]
    qpos0: np.ndarray         # [nv]
    act_dof: np.ndarray       # [nu] int
    gear: np.ndarray          # [nu]
    ctrl_lo: np.ndarray       # [nu]
    ctrl_hi: np.ndarray       # [nu]
    ctrl_limited: np.ndarray  # [nu] bool
    gravity: np.ndarray = field(default_factory=lambda: np.zeros(2))
    dt: float = 0.01
    integrator: str = "RK4"

    @property
    def nbody(self) -> int:
        return len(self.parent)

    @property
    def nv(self) -> int:
        return len(self.dof_body)

    def ancestors_dofs(self, b: int) -> list:
        """Dofs that move body b: its own and its ancestors', in dof order."""
        chain = []
        while b >= 0:
            chain.append(b)
            b = int(self.parent[b])
        bodies = set(chain)
        return [k for k in range(self.nv) if int(self.dof_body[k]) in bodies]


def perp(v):
    return np.array([-v[1], v[0]])


def rot(c, s, v):
    return np.array([c * v[0] - s * v[1], s * v[0] + c * v[1]])


@dataclass
class Kinematics:
    p: np.ndarray       # [nb, 2] body origins
    cs: np.ndarray      # [nb, 2] (cos, sin) of body angles
    c: np.ndarray       # [nb, 2] centres of mass
    cdot: np.ndarray    # [nb, 2] their velocities
    A: np.ndarray       # [nv, 2] hinge anchors (world) / unused for slides
    Adot: np.ndarray    # [nv, 2] anchor velocities
    U: np.ndarray       # [nv, 2] slide axes (world) / unused for hinges
    wb: np.ndarray      # [nv] angular velocity of the frame before each dof


def kinematics(m: ChainModel, q, qd) -> Kinematics:
    nb, nv = m.nbody, m.nv
    P, CS, V, W = np.zeros((nb, 2)), np.zeros((nb, 2)), np.zeros((nb, 2)), np.zeros(nb)
    A, Ad, U, wb = np.zeros((nv, 2)), np.zeros((nv, 2)), np.zeros((nv, 2)), np.zeros(nv)
    C, Cd = np.zeros((nb, 2)), np.zeros((nb, 2))
    for b in range(nb):
        pa = int(m.parent[b])
        if pa < 0:
            p, (c, s), v, w = np.zeros(2), (1.0, 0.0), np.zeros(2), 0.0
        else:
            p, (c, s), v, w = P[pa].copy(), (CS[pa, 0], CS[pa, 1]), V[pa].copy(), W[pa]
        if not m.free[b]:
            off = rot(c, s, m.pos[b])
            p = p + off
            v = v + w * perp(off)
            ca, sa = np.cos(m.angle[b]), np.sin(m.angle[b])
            c, s = c * ca - s * sa, s * ca + c * sa
        for k in range(nv):
            if int(m.dof_body[k]) != b:
                continue
            wb[k] = w
            if m.dof_type[k] == HINGE:
                a = rot(c, s, m.dof_anchor[k])
                A[k] = p + a
                Ad[k] = v + w * perp(a)
                ang = m.dof_sign[k] * q[k]
                ck, sk = np.cos(ang), np.sin(ang)
                c, s = c * ck - s * sk, s * ck + c * sk
                w = w + m.dof_sign[k] * qd[k]
                an = rot(c, s, m.dof_anchor[k])
                p = A[k] - an
                v = Ad[k] - w * perp(an)
            else:
                U[k] = rot(c, s, m.dof_axis[k])
                d = U[k] * q[k]
                p = p + d
                v = v + w * perp(d) + U[k] * qd[k]
        P[b], CS[b], V[b], W[b] = p, (c, s), v, w
        l = rot(c, s, m.com[b])
        C[b] = p + l
        Cd[b] = v + w * perp(l)
    return Kinematics(P, CS, C, Cd, A, Ad, U, wb)


def jacobians(m: ChainModel, K: Kinematics):
    """Centre-of-mass Jacobians J [nb, 2, nv] and angle Jacobians Jphi [nb, nv]."""
    nb, nv = m.nbody, m.nv
    J, Jp = np.zeros((nb, 2, nv)), np.zeros((nb, nv))
    for b in range(nb):
        for k in m.ancestors_dofs(b):
            if m.dof_type[k] == HINGE:
                J[b, :, k] = m.dof_sign[k] * perp(K.c[b] - K.A[k])
                Jp[b, k] = m.dof_sign[k]
            else:
                J[b, :, k] = K.U[k]
    return J, Jp


def bias_accel(m: ChainModel, K: Kinematics, qd):
    """Centre-of-mass accelerations at zero joint acceleration [nb, 2]."""
    a = np.zeros((m.nbody, 2))
    for b in range(m.nbody):
        for k in m.ancestors_dofs(b):
            if m.dof_type[k] == HINGE:
                a[b] += qd[k] * m.dof_sign[k] * perp(K.cdot[b] - K.Adot[k])
            else:
                a[b] += qd[k] * K.wb[k] * perp(K.U[k])
    return a


def mass_matrix(m: ChainModel, q) -> np.ndarray:
    K = kinematics(m, q, np.zeros(m.nv))
    J, Jp = jacobians(m, K)
    M = np.diag(np.asarray(m.armature, np.float64)).copy()
    for b in range(m.nbody):
        M += m.mass[b] * J[b].T @ J[b] + m.inertia[b] * np.outer(Jp[b], Jp[b])
    return M


def xfrc_generalized(m: ChainModel, J, Jp, xfrc) -> np.ndarray:
    """J^T of the planar spatial loads xfrc [nb, 6] (force at the centre of
    mass, torque): sum_b J_b^T (fx, fy) + Jphi_b tz."""
    f = np.zeros(m.nv)
    for b in range(m.nbody):
        f += J[b].T @ xfrc[b, 0:2] + Jp[b] * xfrc[b, 5]
    return f


def actuator_force(m: ChainModel, ctrl) -> np.ndarray:
    f = np.zeros(m.nv)
    for a in range(len(m.act_dof)):
        u = float(ctrl[a])
        if m.ctrl_limited[a]:
            u = min(max(u, float(m.ctrl_lo[a])), float(m.ctrl_hi[a]))
        f[int(m.act_dof[a])] += m.gear[a] * u
    return f


def forward(m: ChainModel, q, qd, ctrl, xfrc, qfrc_applied, implicit_dt: float = 0.0):
    """Joint accelerations (and the pieces) at (q, qd)."""
    K = kinematics(m, q, qd)
    J, Jp = jacobians(m, K)
    a = bias_accel(m, K, qd)
    M = np.diag(np.asarray(m.armature, np.float64)).copy()
    bias = np.zeros(m.nv)
    for b in range(m.nbody):
        M += m.mass[b] * J[b].T @ J[b] + m.inertia[b] * np.outer(Jp[b], Jp[b])
        bias += m.mass[b] * J[b].T @ (a[b] - m.gravity)
    passive = -m.damping * qd - m.stiffness * (q - m.springref)
    f = passive + actuator_force(m, ctrl) + xfrc_generalized(m, J, Jp, xfrc) + qfrc_applied - bias
    if implicit_dt:
        M = M + implicit_dt * np.diag(m.damping)
    return np.linalg.solve(M, f), M, bias, J, Jp


def step(m: ChainModel, q, qd, ctrl, xfrc, applied_factor: float = 1.0, dt: Optional[float] = None):
    """One engine step from (q, qd): returns (q, qd, qfrc_applied)."""
    q = np.asarray(q, np.float64)
    qd = np.asarray(qd, np.float64)
    xfrc = np.asarray(xfrc, np.float64).reshape(m.nbody, 6)
    h = m.dt if dt is None else float(dt)
    K = kinematics(m, q, qd)
    J, Jp = jacobians(m, K)
    qfrc_applied = applied_factor * xfrc_generalized(m, J, Jp, xfrc)
    if m.integrator.upper() == "RK4":
        def f(qq, vv):
            return vv, forward(m, qq, vv, ctrl, xfrc, qfrc_applied)[0]

        k1q, k1v = f(q, qd)
        k2q, k2v = f(q + 0.5 * h * k1q, qd + 0.5 * h * k1v)
        k3q, k3v = f(q + 0.5 * h * k2q, qd + 0.5 * h * k2v)
        k4q, k4v = f(q + h * k3q, qd + h * k3v)
        qn = q + h * (k1q / 6.0 + k2q / 3.0 + k3q / 3.0 + k4q / 6.0)
        vn = qd + h * (k1v / 6.0 + k2v / 3.0 + k3v / 3.0 + k4v / 6.0)
        return qn, vn, qfrc_applied
    acc = forward(m, q, qd, ctrl, xfrc, qfrc_applied, implicit_dt=h)[0]
    vn = qd + h * acc
    return q + h * vn, vn, qfrc_applied


def poses(m: ChainModel, q):
    """Body origins xpos [nb, 3], centres of mass xipos [nb, 3] and body
    rotations xquat [nb, 4] (w, x, y, z) about z."""
    K = kinematics(m, q, np.zeros(m.nv))
    nb = m.nbody
    xpos, xipos, xquat = np.zeros((nb, 3)), np.zeros((nb, 3)), np.zeros((nb, 4))
    for b in range(nb):
        xpos[b] = (K.p[b, 0], K.p[b, 1], m.posz[b])
        xipos[b] = (K.c[b, 0], K.c[b, 1], m.comz[b])
        phi = np.arctan2(K.cs[b, 1], K.cs[b, 0])
        xquat[b] = (np.cos(0.5 * phi), 0.0, 0.0, np.sin(0.5 * phi))
    return xpos, xipos, xquat


def energy(m: ChainModel, q, qd) -> float:
    """Kinetic energy 1/2 qd^T M qd (armature included)."""
    return 0.5 * float(qd @ mass_matrix(m, q) @ qd)

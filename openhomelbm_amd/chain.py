"""Planar articulated chains on the device, and their two-way coupling.

The reference's coupled environments step MuJoCo-Warp once per lattice step,
with the fluid loads as xfrc_applied and, accumulated through J^T, as
qfrc_applied (envs/lbm/lbm_fluid_env.py:441-533), on MJCF chains such as
envs/lbm/eel/eel_2d.xml. MuJoCo does not run on this stack; `Chain` steps
the same planar model on the device (hl_chain_step: joint-space inertia and
bias forces of the tree, joint damping / stiffness / armature, motors, RK4 or
semi-implicit Euler, FP64), and `ChainWorlds` runs the coupled loop of
_simulation_step with it: chain poses into the lattices' pose tables and
capsule masks, a lattice step with its FP64 force rows, rows into xfrc, a
chain step; all on the device.

`chain_from_mjcf` reads the model from an MJCF file: option (timestep,
gravity, integrator), compiler angle units, default classes (joint damping,
stiffness, armature, springref, ref; motor gear and ctrlrange; geom density
and mass), bodies with their joints, geoms (mass, centre of mass and inertia
from the shapes at uniform density, as MuJoCo's inertiafromgeom) or an
inertial element, and motor actuators. Planar models only: hinge axes along
z, slide axes in the plane, body rotations about z, a free joint read as the
three planar dofs of its body. Joint limits are recorded but not enforced,
and there are no contacts.
"""
from __future__ import annotations

import ctypes as C
import os
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _native as N
from .lattice import Lattice, _stream_ptr

SLIDE, HINGE = N.HL_CHAIN_SLIDE, N.HL_CHAIN_HINGE


@dataclass
class ChainModel:
    """A planar kinematic tree (the fields of struct hl_chain_model): bodies
    in tree order, each body's dofs consecutive in joint order."""

    parent: np.ndarray
    free: np.ndarray
    pos: np.ndarray
    angle: np.ndarray
    com: np.ndarray
    mass: np.ndarray
    inertia: np.ndarray
    posz: np.ndarray
    comz: np.ndarray
    dof_body: np.ndarray
    dof_type: np.ndarray
    dof_axis: np.ndarray
    dof_anchor: np.ndarray
    dof_sign: np.ndarray
    damping: np.ndarray
    stiffness: np.ndarray
    springref: np.ndarray
    armature: np.ndarray
    qpos0: np.ndarray
    act_dof: np.ndarray
    gear: np.ndarray
    ctrl_lo: np.ndarray
    ctrl_hi: np.ndarray
    ctrl_limited: np.ndarray
    gravity: np.ndarray = field(default_factory=lambda: np.zeros(2))
    dt: float = 0.01
    integrator: str = "RK4"
    applied_factor: float = 1.0
    body_names: List[str] = field(default_factory=list)
    joint_names: List[str] = field(default_factory=list)   # one per dof (free joints: name_x, name_y, name_rz)
    act_names: List[str] = field(default_factory=list)
    dof_range: Optional[np.ndarray] = None                # [nv, 2] radians / lengths; +-inf when unlimited
    geoms: List[dict] = field(default_factory=list)       # per geom: body, type, size, pos (body frame), mat

    @property
    def nbody(self) -> int:
        return len(self.parent)

    @property
    def nv(self) -> int:
        return len(self.dof_body)

    @property
    def nact(self) -> int:
        return len(self.act_dof)

    def body_id(self, name: str) -> int:
        return self.body_names.index(name)

    def to_struct(self) -> N.HLChainModel:
        s = N.HLChainModel()
        nb, nv, nu = self.nbody, self.nv, self.nact
        if not (1 <= nb <= N.HL_CHAIN_MAX_BODIES and 1 <= nv <= N.HL_CHAIN_MAX_DOFS and nu <= N.HL_CHAIN_MAX_ACTS):
            raise ValueError(f"a chain holds 1..16 bodies, 1..16 dofs and 0..16 motors (got {nb}, {nv}, {nu})")
        s.nbody, s.ndof, s.nact = nb, nv, nu
        s.integrator = N.HL_CHAIN_RK4 if self.integrator.upper() == "RK4" else N.HL_CHAIN_EULER
        for b in range(nb):
            s.parent[b], s.free_body[b] = int(self.parent[b]), int(bool(self.free[b]))
            s.pos[b][0], s.pos[b][1] = float(self.pos[b][0]), float(self.pos[b][1])
            s.angle[b] = float(self.angle[b])
            s.com[b][0], s.com[b][1] = float(self.com[b][0]), float(self.com[b][1])
            s.mass[b], s.inertia[b] = float(self.mass[b]), float(self.inertia[b])
            s.posz[b], s.comz[b] = float(self.posz[b]), float(self.comz[b])
        for k in range(nv):
            s.dof_body[k], s.dof_type[k] = int(self.dof_body[k]), int(self.dof_type[k])
            s.dof_axis[k][0], s.dof_axis[k][1] = float(self.dof_axis[k][0]), float(self.dof_axis[k][1])
            s.dof_anchor[k][0], s.dof_anchor[k][1] = float(self.dof_anchor[k][0]), float(self.dof_anchor[k][1])
            s.dof_sign[k] = float(self.dof_sign[k])
            s.damping[k], s.stiffness[k] = float(self.damping[k]), float(self.stiffness[k])
            s.springref[k], s.armature[k] = float(self.springref[k]), float(self.armature[k])
        for a in range(nu):
            s.act_dof[a], s.ctrl_limited[a] = int(self.act_dof[a]), int(bool(self.ctrl_limited[a]))
            s.gear[a], s.ctrl_lo[a], s.ctrl_hi[a] = float(self.gear[a]), float(self.ctrl_lo[a]), float(self.ctrl_hi[a])
        s.gravity[0], s.gravity[1] = float(self.gravity[0]), float(self.gravity[1])
        s.dt, s.applied_factor = float(self.dt), float(self.applied_factor)
        return s


# ------------------------------------------------------------ geometry

def shape_volume(gtype: str, size: np.ndarray) -> float:
    r = float(size[0])
    if gtype == "sphere":
        return 4.0 / 3.0 * np.pi * r ** 3
    if gtype == "capsule":
        return np.pi * r * r * 2.0 * size[1] + 4.0 / 3.0 * np.pi * r ** 3
    if gtype == "cylinder":
        return np.pi * r * r * 2.0 * size[1]
    if gtype == "box":
        return 8.0 * size[0] * size[1] * size[2]
    if gtype == "ellipsoid":
        return 4.0 / 3.0 * np.pi * size[0] * size[1] * size[2]
    raise NotImplementedError(f"mass properties of {gtype} geoms")


def shape_inertia(gtype: str, size: np.ndarray, mass: float) -> np.ndarray:
    """Principal moments about the centre, in the geom frame (capsules and
    cylinders along z; size = MuJoCo's: radius, half length / half extents)."""
    r = float(size[0])
    if gtype == "sphere":
        i = 0.4 * mass * r * r
        return np.array([i, i, i])
    if gtype == "cylinder":
        h = 2.0 * size[1]
        ip = mass * (3.0 * r * r + h * h) / 12.0
        return np.array([ip, ip, 0.5 * mass * r * r])
    if gtype == "capsule":
        # a cylinder of length h and two hemispheres whose centres of mass sit
        # 3r/8 beyond its ends: I_perp = m_c (3 r^2 + h^2) / 12 + m_s (2 r^2 / 5
        # + h^2 / 4 + 3 h r / 8), I_axial = m_c r^2 / 2 + m_s 2 r^2 / 5
        h = 2.0 * size[1]
        vs, vc = 4.0 / 3.0 * np.pi * r ** 3, np.pi * r * r * h
        ms = mass * vs / (vs + vc)
        mc = mass - ms
        ip = mc * (3.0 * r * r + h * h) / 12.0 + ms * (0.4 * r * r + 0.25 * h * h + 0.375 * h * r)
        ia = 0.5 * mc * r * r + 0.4 * ms * r * r
        return np.array([ip, ip, ia])
    if gtype == "box":
        a, b, c = (float(v) for v in size[:3])
        return mass / 3.0 * np.array([b * b + c * c, a * a + c * c, a * a + b * b])
    if gtype == "ellipsoid":
        a, b, c = (float(v) for v in size[:3])
        return mass / 5.0 * np.array([b * b + c * c, a * a + c * c, a * a + b * b])
    raise NotImplementedError(f"mass properties of {gtype} geoms")


# ------------------------------------------------------------ MJCF

def _f(s: Optional[str], default: Optional[str] = None) -> Optional[np.ndarray]:
    if s is None:
        s = default
    if s is None:
        return None
    return np.array([float(x) for x in s.split()], np.float64)


def _rotz(a: float) -> np.ndarray:
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


class _Defaults:
    """MJCF default classes: attributes per element tag, inherited from the
    enclosing class ('main' at the top)."""

    def __init__(self, root):
        self.classes: Dict[str, Dict[str, dict]] = {}
        top = root.find("default")
        if top is not None:
            self._walk(top, "main", {})
        self.classes.setdefault("main", {})

    def _walk(self, elem, name, inherited):
        own = {tag: dict(attrs) for tag, attrs in inherited.items()}
        for child in elem:
            if child.tag == "default":
                continue
            own.setdefault(child.tag, {}).update(child.attrib)
        self.classes[name] = own
        for child in elem.findall("default"):
            self._walk(child, child.get("class", name), own)

    def attrs(self, elem, cls: str, tag: Optional[str] = None) -> dict:
        c = elem.get("class", cls)
        tags = (tag or elem.tag,) if tag is None or tag == elem.tag else (tag, elem.tag)
        out = {}
        for t in tags:
            out.update(self.classes.get(c, self.classes["main"]).get(t, {}))
        out.update(elem.attrib)
        return out


def chain_from_mjcf(path: str, applied_factor: float = 1.0) -> ChainModel:
    """The planar chain of an MJCF file (see the module docstring)."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"MJCF file not found: {path}")
    root = ET.parse(path).getroot()
    comp = root.find("compiler")
    deg = comp is None or comp.get("angle", "degree") == "degree"
    to_rad = (np.pi / 180.0) if deg else 1.0
    opt = root.find("option")
    dt = float(opt.get("timestep", "0.002")) if opt is not None else 0.002
    grav = _f(opt.get("gravity") if opt is not None else None, "0 0 -9.81")
    integ = (opt.get("integrator", "Euler") if opt is not None else "Euler")
    if integ.upper() not in ("RK4", "EULER"):
        raise NotImplementedError(f"integrator {integ} (RK4 and Euler are supported)")
    D = _Defaults(root)

    def orient(a: dict) -> np.ndarray:
        if "quat" in a:
            from .mjcf import quat_to_matrix
            return quat_to_matrix(_f(a["quat"]))
        if "euler" in a:
            e = _f(a["euler"]) * to_rad
            Rx = np.array([[1, 0, 0], [0, np.cos(e[0]), -np.sin(e[0])], [0, np.sin(e[0]), np.cos(e[0])]])
            Ry = np.array([[np.cos(e[1]), 0, np.sin(e[1])], [0, 1, 0], [-np.sin(e[1]), 0, np.cos(e[1])]])
            return Rx @ Ry @ _rotz(e[2])  # MuJoCo's default eulerseq 'xyz': axes moving with the frame
        if "axisangle" in a:
            v = _f(a["axisangle"])
            ax = v[:3] / np.linalg.norm(v[:3])
            ang = v[3] * to_rad
            K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
            return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
        return np.eye(3)

    def planar_angle(R: np.ndarray, what: str) -> float:
        if abs(R[2, 2] - 1.0) > 1e-9 or abs(R[0, 2]) > 1e-9 or abs(R[1, 2]) > 1e-9:
            raise NotImplementedError(f"{what} is not a rotation about z (planar chains only)")
        return float(np.arctan2(R[1, 0], R[0, 0]))

    parent, free, pos, angle, posz = [], [], [], [], []
    mass, com, comz, inertia, names = [], [], [], [], []
    dof = {k: [] for k in ("body", "type", "axis", "anchor", "sign", "damping", "stiffness", "springref",
                           "armature", "qpos0", "range", "name")}
    geoms: List[dict] = []
    world_R = {-1: np.eye(3)}
    world_P = {-1: np.zeros(3)}

    def add_dof(b, typ, axis, anchor, sign, ja, q0, name, rng):
        dof["body"].append(b)
        dof["type"].append(typ)
        dof["axis"].append(axis)
        dof["anchor"].append(anchor)
        dof["sign"].append(sign)
        dof["damping"].append(float(ja.get("damping", "0")))
        dof["stiffness"].append(float(ja.get("stiffness", "0")))
        scale = to_rad if typ == HINGE else 1.0
        dof["springref"].append(float(ja.get("springref", "0")) * scale)
        dof["armature"].append(float(ja.get("armature", "0")))
        dof["qpos0"].append(q0)
        dof["name"].append(name)
        dof["range"].append(rng)

    def body_mass(belem, cls, bid, Rw, Pw):
        inertial = belem.find("inertial")
        gl = []
        for g in belem.findall("geom"):
            a = D.attrs(g, cls)
            gtype = a.get("type", "sphere")
            if gtype == "plane":
                continue
            size = _f(a.get("size"), "0.5 0.5 0.5")
            if "fromto" in a and gtype in ("capsule", "cylinder"):
                ft = _f(a["fromto"])
                p1, p2 = ft[:3], ft[3:]
                L = float(np.linalg.norm(p2 - p1))
                from .mjcf import _align_z
                d = (p2 - p1) / L if L > 1e-12 else np.array([0.0, 0.0, 1.0])
                gp, gR = 0.5 * (p1 + p2), _align_z(d)
                size = np.array([size[0], 0.5 * L, 0.0])
            else:
                gp, gR = _f(a.get("pos"), "0 0 0"), orient(a)
                size = np.concatenate([size, np.zeros(3)])[:3]
                if gtype in ("capsule", "cylinder") and len(_f(a.get("size"), "0.5")) == 1:
                    size[1] = size[0]
                if gtype == "sphere":
                    size[1:] = 0.0
            if gtype == "mesh":
                raise NotImplementedError("mesh geoms in a chain (give the body an inertial element)")
            vol = shape_volume(gtype, size)
            m = float(a["mass"]) if "mass" in a else float(a.get("density", "1000")) * vol
            gl.append((gtype, size, gp, gR, m))
            geoms.append({"body": bid, "type": gtype, "size": size.copy(), "pos": gp.copy(), "mat": gR.copy(),
                          "name": g.get("name", f"geom_{len(geoms)}")})
        if inertial is not None:
            ia = inertial.attrib
            m = float(ia.get("mass", "0"))
            c = _f(ia.get("pos"), "0 0 0")
            if "fullinertia" in ia:
                fi = _f(ia["fullinertia"])
                Ib = np.array([[fi[0], fi[3], fi[4]], [fi[3], fi[1], fi[5]], [fi[4], fi[5], fi[2]]])
            else:
                Ri = orient(ia)
                Ib = Ri @ np.diag(_f(ia.get("diaginertia"), "0 0 0")) @ Ri.T
            return m, c, Ib
        mt = sum(g[4] for g in gl)
        if mt <= 0.0:
            return 0.0, np.zeros(3), np.zeros((3, 3))
        c = sum(g[4] * g[2] for g in gl) / mt
        Ib = np.zeros((3, 3))
        for gtype, size, gp, gR, m in gl:
            Ig = gR @ np.diag(shape_inertia(gtype, size, m)) @ gR.T
            d = gp - c
            Ib += Ig + m * (np.dot(d, d) * np.eye(3) - np.outer(d, d))
        return mt, c, Ib

    def walk(elem, pid, cls):
        for belem in elem.findall("body"):
            bcls = belem.get("childclass", cls)
            bid = len(parent)
            name = belem.get("name", f"body_{bid}")
            lp = _f(belem.get("pos"), "0 0 0")
            lR = orient(belem.attrib)
            Rw = world_R[pid] @ lR
            Pw = world_P[pid] + world_R[pid] @ lp
            world_R[bid], world_P[bid] = Rw, Pw
            a_loc = planar_angle(lR, f"body {name!r}'s orientation")
            a_w = planar_angle(Rw, f"body {name!r}'s orientation")
            joints = [D.attrs(j, bcls) | {"_tag": "joint"} for j in belem.findall("joint")]
            joints += [D.attrs(j, bcls, "joint") | {"type": "free", "_tag": "freejoint"}
                       for j in belem.findall("freejoint")]
            is_free = any(j.get("type", "hinge") == "free" for j in joints)
            if is_free and len(joints) > 1:
                raise NotImplementedError(f"body {name!r}: a free joint with other joints")
            parent.append(pid)
            free.append(is_free)
            pos.append(lp[:2])
            angle.append(a_loc)
            posz.append(float(Pw[2]))
            names.append(name)
            m, c, Ib = body_mass(belem, bcls, bid, Rw, Pw)
            tr = max(np.trace(Ib), 1e-300)
            if abs(Ib[0, 2]) > 1e-9 * tr or abs(Ib[1, 2]) > 1e-9 * tr:
                raise NotImplementedError(f"body {name!r}: inertia couples z with x or y (not planar)")
            mass.append(m)
            com.append(c[:2])
            comz.append(float(Pw[2] + (Rw @ c)[2]))
            inertia.append(float(Ib[2, 2]))
            for j in joints:
                jtype = j.get("type", "hinge")
                jname = j.get("name", f"joint_{len(dof['name'])}")
                lim = j.get("limited", "auto")
                rng = _f(j.get("range"), "0 0")
                limited = lim == "true" or (lim == "auto" and "range" in j)
                if jtype == "free":
                    add_dof(bid, SLIDE, np.array([1.0, 0.0]), np.zeros(2), 1.0, j, float(Pw[0]), jname + "_x",
                            (-np.inf, np.inf))
                    add_dof(bid, SLIDE, np.array([0.0, 1.0]), np.zeros(2), 1.0, j, float(Pw[1]), jname + "_y",
                            (-np.inf, np.inf))
                    add_dof(bid, HINGE, np.zeros(2), np.zeros(2), 1.0, j, a_w, jname + "_rz", (-np.inf, np.inf))
                elif jtype == "hinge":
                    ax = _f(j.get("axis"), "0 0 1")
                    ax = ax / np.linalg.norm(ax)
                    if abs(abs(ax[2]) - 1.0) > 1e-9:
                        raise NotImplementedError(f"joint {jname!r}: hinge axis not along z (planar chains only)")
                    anchor = _f(j.get("pos"), "0 0 0")[:2]
                    r = (rng[0] * to_rad, rng[1] * to_rad) if limited else (-np.inf, np.inf)
                    add_dof(bid, HINGE, np.zeros(2), anchor, float(np.sign(ax[2])), j,
                            float(j.get("ref", "0")) * to_rad, jname, r)
                elif jtype == "slide":
                    ax = _f(j.get("axis"), "0 0 1")
                    ax = ax / np.linalg.norm(ax)
                    if abs(ax[2]) > 1e-9:
                        raise NotImplementedError(f"joint {jname!r}: slide axis leaves the plane")
                    r = (rng[0], rng[1]) if limited else (-np.inf, np.inf)
                    add_dof(bid, SLIDE, ax[:2], np.zeros(2), 1.0, j, float(j.get("ref", "0")), jname, r)
                else:
                    raise NotImplementedError(f"joint {jname!r}: {jtype} joints (free, hinge and slide only)")
            walk(belem, bid, bcls)

    wb = root.find("worldbody")
    if wb is not None:
        walk(wb, -1, "main")
    if abs(grav[2]) > 0.0 and any(free):
        raise NotImplementedError("gravity along z would take a free body out of the plane")
    if not parent:
        raise ValueError(f"{path}: no bodies")
    if not dof["body"]:
        raise ValueError(f"{path}: no joints (nothing moves)")
    # motors
    act_dof, gear, lo, hi, lim, anames = [], [], [], [], [], []
    act = root.find("actuator")
    if act is not None:
        for e in act:
            if e.tag != "motor":
                raise NotImplementedError(f"actuator <{e.tag}> (motors only)")
            a = D.attrs(e, "main")
            jn = a.get("joint")
            cand = [k for k, n in enumerate(dof["name"]) if n == jn]
            if len(cand) != 1:
                raise NotImplementedError(f"motor on {jn!r}: hinge or slide joints only")
            act_dof.append(cand[0])
            gear.append(float(_f(a.get("gear"), "1")[0]))
            cr = _f(a.get("ctrlrange"), "0 0")
            cl = a.get("ctrllimited", "auto")
            lim.append(cl == "true" or (cl == "auto" and "ctrlrange" in a))
            lo.append(float(cr[0]))
            hi.append(float(cr[1]))
            anames.append(e.get("name", f"motor_{len(anames)}"))
    nb = len(parent)
    return ChainModel(
        parent=np.array(parent, np.int32), free=np.array(free, bool), pos=np.array(pos).reshape(nb, 2),
        angle=np.array(angle), com=np.array(com).reshape(nb, 2), mass=np.array(mass),
        inertia=np.array(inertia), posz=np.array(posz), comz=np.array(comz),
        dof_body=np.array(dof["body"], np.int32), dof_type=np.array(dof["type"], np.int32),
        dof_axis=np.array(dof["axis"]).reshape(-1, 2), dof_anchor=np.array(dof["anchor"]).reshape(-1, 2),
        dof_sign=np.array(dof["sign"]), damping=np.array(dof["damping"]), stiffness=np.array(dof["stiffness"]),
        springref=np.array(dof["springref"]), armature=np.array(dof["armature"]), qpos0=np.array(dof["qpos0"]),
        act_dof=np.array(act_dof, np.int32), gear=np.array(gear), ctrl_lo=np.array(lo), ctrl_hi=np.array(hi),
        ctrl_limited=np.array(lim, bool), gravity=grav[:2].copy(), dt=dt, integrator=integ.upper(),
        applied_factor=float(applied_factor), body_names=names, joint_names=list(dof["name"]),
        act_names=anames, dof_range=np.array(dof["range"], np.float64).reshape(-1, 2), geoms=geoms)


# ------------------------------------------------------------ device chain

class Chain:
    """nworld copies of a chain on the device: qpos / qvel FP64 [nworld, nv],
    ctrl FP32 [nworld, nact], xfrc FP32 [nworld, nbody, 6] (MuJoCo's
    xfrc_applied layout: force at the centre of mass, torque), and after
    each step the poses xpos / xipos [nworld, nbody, 3], xquat [nworld,
    nbody, 4] and qfrc_applied [nworld, nv]."""

    def __init__(self, model: ChainModel, nworld: int = 1, device=None):
        self.model = model
        self.nworld = int(nworld)
        dev = torch.device(device or "cuda")
        self.device = dev if dev.index is not None else torch.device("cuda", torch.cuda.current_device())
        self.lib = N.load()
        h = C.c_void_p()
        st = model.to_struct()
        N.check(self.lib.hl_chain_create(C.byref(st), C.byref(h)), "hl_chain_create")
        self._h = h
        nb, nv, nu = model.nbody, model.nv, model.nact
        dev = self.device
        self.qpos = torch.tensor(np.tile(model.qpos0, (self.nworld, 1)), dtype=torch.float64, device=dev)
        self.qvel = torch.zeros((self.nworld, nv), dtype=torch.float64, device=dev)
        self.ctrl = torch.zeros((self.nworld, max(nu, 1)), dtype=torch.float32, device=dev)
        self.xfrc = torch.zeros((self.nworld, nb, 6), dtype=torch.float32, device=dev)
        self.qfrc_applied = torch.zeros((self.nworld, nv), dtype=torch.float64, device=dev)
        self.xpos = torch.zeros((self.nworld, nb, 3), dtype=torch.float32, device=dev)
        self.xipos = torch.zeros_like(self.xpos)
        self.xquat = torch.zeros((self.nworld, nb, 4), dtype=torch.float32, device=dev)
        self.step(0)  # poses of the initial state

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            N.check(self.lib.hl_chain_destroy(self._h), "hl_chain_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, nsteps: int = 1, ctrl: Optional[torch.Tensor] = None, xfrc: Optional[torch.Tensor] = None,
             stream=None) -> None:
        """nsteps engine steps with ctrl / xfrc (default: the chain's own
        buffers) held over them."""
        c = self.ctrl if ctrl is None else ctrl
        x = self.xfrc if xfrc is None else xfrc
        for t, shape in ((c, (self.nworld, max(self.model.nact, 1))), (x, (self.nworld, self.model.nbody, 6))):
            if t.dtype != torch.float32 or not t.is_contiguous() or t.device != self.device or t.shape[0] != shape[0] \
                    or t.numel() < int(np.prod(shape)):
                raise ValueError(f"ctrl / xfrc must be contiguous FP32 device tensors of shape {shape}")
        N.check(self.lib.hl_chain_step(self._h, self.nworld, self.qpos.data_ptr(), self.qvel.data_ptr(),
                                       c.data_ptr() if self.model.nact else None, x.data_ptr(), int(nsteps),
                                       self.qfrc_applied.data_ptr(), self.xpos.data_ptr(), self.xipos.data_ptr(),
                                       self.xquat.data_ptr(), _stream_ptr(stream)), "hl_chain_step")

    def limits_exceeded(self) -> torch.Tensor:
        """[nworld] bool: some joint outside its range (limits are not enforced)."""
        r = self.model.dof_range
        if r is None:
            return torch.zeros(self.nworld, dtype=torch.bool, device=self.device)
        lo = torch.tensor(r[:, 0], dtype=torch.float64, device=self.device)
        hi = torch.tensor(r[:, 1], dtype=torch.float64, device=self.device)
        return ((self.qpos < lo) | (self.qpos > hi)).any(dim=1)


def capsule_shapes(model: ChainModel, bodies: Sequence[int]) -> np.ndarray:
    """Body-frame capsule segments [len(bodies), 7] (a, b, radius) of the
    given bodies, one capsule or sphere geom each."""
    out = np.zeros((len(bodies), 7))
    for i, b in enumerate(bodies):
        gs = [g for g in model.geoms if g["body"] == b and g["type"] in ("capsule", "sphere")]
        if len(gs) != 1:
            raise NotImplementedError(f"body {model.body_names[b]!r}: one capsule or sphere geom per coupled body")
        g = gs[0]
        half = g["mat"] @ np.array([0.0, 0.0, g["size"][1] if g["type"] == "capsule" else 0.0])
        out[i, 0:3] = g["pos"] - half
        out[i, 3:6] = g["pos"] + half
        out[i, 6] = g["size"][0]
    return out


class ChainWorlds:
    """The reference's coupled sub-step loop (lbm_fluid_env.py:441-533) for
    nworld D2Q9 lattices and a chain per world, on the device. Per step:
    chain poses -> the lattices' pose tables (hl_couple_poses: centre of
    mass and angle, the previous pose kept as 'last') and posed capsules ->
    masks inside `box`; one lattice step with its FP64 force rows; rows ->
    xfrc (hl_couple_forces, the reference's force / scale, torque / scale);
    one chain step with ctrl. bodies: the chain bodies coupled to the fluid,
    body k as solid solid_ids[k]; mapping lattice = (engine -
    engine_origin) * scale + lattice_origin per body."""

    def __init__(self, lattices: Sequence[Lattice], chain: Chain, bodies: Sequence[int], solid_ids: Sequence[int],
                 box, engine_origin=None, lattice_origin=None, scale=None):
        from .coupling import BodyMap, _tables
        if len(lattices) != chain.nworld:
            raise ValueError("one lattice per world")
        if any(L.lattice != N.HL_D2Q9 for L in lattices):
            raise NotImplementedError("ChainWorlds couples D2Q9 lattices (planar chains)")
        self.lattices, self.chain = list(lattices), chain
        self.bodies = [int(b) for b in bodies]
        dev = chain.device
        self.bmap = BodyMap(solid_ids, engine_origin, lattice_origin, scale, body_ids=self.bodies, device=dev)
        self.shapes = torch.tensor(capsule_shapes(chain.model, self.bodies), dtype=torch.float32, device=dev)
        box = tuple(int(v) for v in box)
        self.box = box + (0, 1) if len(box) == 4 else box  # x0, x1, y0, y1 (, z0, z1)
        self.nsolid = self.lattices[0].nsolid
        self.rows = torch.zeros((chain.nworld, 1, self.nsolid, 6), dtype=torch.float64, device=dev)
        self.init = torch.zeros((chain.nworld, self.nsolid), dtype=torch.int32, device=dev)
        self.tables = _tables(self.lattices)
        nb = len(self.bodies)
        self.caps = torch.empty((chain.nworld, nb, 8), dtype=torch.float32, device=dev)
        self._sel = torch.tensor(self.bodies, dtype=torch.long, device=dev)

    def step(self, nsteps: int = 1, ctrl: Optional[torch.Tensor] = None, stream=None) -> None:
        from .coupling import forces_to_engine, poses_to_lattices
        ch = self.chain
        lib = N.load()
        for _ in range(nsteps):
            xpos = ch.xpos.index_select(1, self._sel).contiguous()
            xipos = ch.xipos.index_select(1, self._sel).contiguous()
            xquat = ch.xquat.index_select(1, self._sel).contiguous()
            bm = self.bmap
            N.check(lib.hl_body_capsules(ch.nworld, len(self.bodies), xpos.data_ptr(), 3, xquat.data_ptr(), 4,
                                         self.shapes.data_ptr(), bm.solid_ids.data_ptr(),
                                         bm.engine_origin.data_ptr(), bm.lattice_origin.data_ptr(),
                                         bm.scale.data_ptr(), self.caps.data_ptr(), _stream_ptr(stream)),
                    "hl_body_capsules")
            for w, lat in enumerate(self.lattices):
                lat.voxelize_device(self.caps[w], self.box, stream=stream)
            poses_to_lattices(self.lattices, bm, xipos, xquat, self.init, 3, 4, tables=self.tables, stream=stream)
            for w, lat in enumerate(self.lattices):
                lat.step(1, N.HL_KERNEL_AUTO, force_history=self.rows[w], stream=stream)
            ch.xfrc.zero_()
            forces_to_engine(N.HL_D2Q9, self.rows, ch.nworld, self.nsolid, bm, N.HL_COUPLE_SCALED, xfrc=ch.xfrc,
                             stream=stream)
            ch.step(1, ctrl=ctrl, stream=stream)

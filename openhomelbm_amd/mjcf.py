"""MJCF geometry for 3D bodies: the reference's MJCF reader
(envs/lbm3d/mjcf_parser.py) without trimesh or MuJoCo.

Same functions, arguments and returned structures: parse_mjcf,
parse_mjcf_to_meshes, parse_mjcf_as_urdf_format, get_body_world_positions.
A geometry is a TriMesh (vertices [n, 3] float64, faces [m, 3] int64, closed,
outward-oriented), which LBM_Solver3D.create_solid_from_mesh takes as it
takes a trimesh.Trimesh.

Geoms (mjcf_parser.py:155-299), each built in its own frame and then moved by
the geom's quat/euler and pos (fromto capsules and cylinders are placed by
their end points instead, as the reference does):
  box       size = half extents; 8 vertices, 12 triangles
  sphere    size = radius; icosphere of 3 subdivisions (642 vertices)
  ellipsoid size = radii; the same icosphere scaled per axis
  cylinder  size = radius, half length (or fromto); 32-sided prism
  capsule   size = radius, half length (or fromto); 16 sections around,
            4 latitude bands per end cap
  mesh      an STL (binary or ASCII) or OBJ file of an asset, with its scale
The sphere, cylinder and capsule tessellations are inscribed in the exact
surfaces (every vertex on them); their sides and band counts follow the
reference's creation calls, the vertex layout is this module's own.

Euler angles are degrees, extrinsic x-y-z, as in the reference (it ignores
the compiler's angle and eulerseq settings); defaults classes and includes
are not expanded, as in the reference.
"""
from __future__ import annotations

import os
import struct
import xml.etree.ElementTree as ET
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np


# ----------------------------------------------------------------- meshes

@dataclass
class TriMesh:
    """Indexed triangle mesh (float64 vertices, int64 faces)."""

    vertices: np.ndarray
    faces: np.ndarray

    def __post_init__(self):
        self.vertices = np.asarray(self.vertices, np.float64).reshape(-1, 3)
        self.faces = np.asarray(self.faces, np.int64).reshape(-1, 3)

    @property
    def volume(self) -> float:
        """Signed volume (positive for a closed outward-oriented mesh)."""
        a, b, c = (self.vertices[self.faces[:, k]] for k in range(3))
        return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)

    @property
    def is_watertight(self) -> bool:
        """Every edge is shared by exactly two faces, once in each direction."""
        e = np.concatenate([self.faces[:, [0, 1]], self.faces[:, [1, 2]], self.faces[:, [2, 0]]])
        fwd = {tuple(x) for x in e.tolist()}
        return len(fwd) == len(e) and all((b, a) in fwd for a, b in fwd)

    def copy(self) -> "TriMesh":
        return TriMesh(self.vertices.copy(), self.faces.copy())

    def transformed(self, R: np.ndarray, t: Sequence[float]) -> "TriMesh":
        return TriMesh(self.vertices @ np.asarray(R, np.float64).T + np.asarray(t, np.float64), self.faces.copy())


def concatenate(meshes: Sequence[TriMesh]) -> TriMesh:
    """One mesh holding every input's triangles (no merging, as
    trimesh.util.concatenate)."""
    vs, fs, n = [], [], 0
    for m in meshes:
        vs.append(m.vertices)
        fs.append(m.faces + n)
        n += len(m.vertices)
    return TriMesh(np.concatenate(vs), np.concatenate(fs))


def box(extents: Sequence[float]) -> TriMesh:
    """Axis-aligned box of the given full extents, centred at the origin."""
    h = 0.5 * np.asarray(extents, np.float64)
    s = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], np.float64)
    f = [[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
         [2, 3, 7], [2, 7, 6], [1, 2, 6], [1, 6, 5], [0, 4, 7], [0, 7, 3]]
    return TriMesh(s * h, f)


def icosphere(subdivisions: int = 3, radius: float = 1.0) -> TriMesh:
    """Icosahedron refined `subdivisions` times (each triangle into four at
    its edge midpoints), every vertex pushed onto the sphere."""
    p = (1.0 + 5.0 ** 0.5) / 2.0
    v = [[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0], [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
         [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    verts = [np.asarray(x, np.float64) / np.linalg.norm(x) for x in v]
    for _ in range(subdivisions):
        mid: Dict[tuple, int] = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                mid[key] = len(verts) - 1
            return mid[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    return TriMesh(np.array(verts) * radius, f)


def cylinder(radius: float, height: float, sections: int = 32) -> TriMesh:
    """Regular `sections`-sided prism along z from -height/2 to height/2,
    its vertices on the circle of the given radius, with fan caps."""
    a = 2.0 * np.pi * np.arange(sections) / sections
    ring = np.stack([radius * np.cos(a), radius * np.sin(a)], 1)
    z0, z1 = -0.5 * height, 0.5 * height
    v = np.concatenate([np.c_[ring, np.full(sections, z0)], np.c_[ring, np.full(sections, z1)],
                        [[0.0, 0.0, z0], [0.0, 0.0, z1]]])
    lo, hi, c0, c1 = 0, sections, 2 * sections, 2 * sections + 1
    f = []
    for k in range(sections):
        j = (k + 1) % sections
        f += [[lo + k, lo + j, hi + j], [lo + k, hi + j, hi + k],  # side
              [c0, lo + j, lo + k], [c1, hi + k, hi + j]]          # caps
    return TriMesh(v, f)


def capsule(height: float, radius: float, count: Sequence[int] = (16, 8)) -> TriMesh:
    """Capsule along z: the cylinder of the given height between two
    hemispherical caps, count[0] sections around and count[1] latitude bands
    over the two caps (count[1] // 2 per cap); vertices on the exact surface."""
    n = int(count[0])
    bands = max(int(count[1]) // 2, 1)
    a = 2.0 * np.pi * np.arange(n) / n
    rings = []  # (ring radius, z) from the bottom pole's ring upwards
    for k in range(1, bands + 1):  # bottom cap, latitude from -90 deg towards 0
        phi = -0.5 * np.pi + 0.5 * np.pi * k / bands
        rings.append((radius * np.cos(phi), -0.5 * height + radius * np.sin(phi)))
    for k in range(0, bands):  # top cap, from 0 towards +90 deg
        phi = 0.5 * np.pi * k / bands
        rings.append((radius * np.cos(phi), 0.5 * height + radius * np.sin(phi)))
    v = [[0.0, 0.0, -0.5 * height - radius]]
    for rr, z in rings:
        v += [[rr * np.cos(t), rr * np.sin(t), z] for t in a]
    v.append([0.0, 0.0, 0.5 * height + radius])
    top = len(v) - 1
    f = []
    for k in range(n):  # bottom pole fan
        f.append([0, 1 + (k + 1) % n, 1 + k])
    for r in range(len(rings) - 1):
        b0, b1 = 1 + r * n, 1 + (r + 1) * n
        for k in range(n):
            j = (k + 1) % n
            f += [[b0 + k, b0 + j, b1 + j], [b0 + k, b1 + j, b1 + k]]
    last = 1 + (len(rings) - 1) * n
    for k in range(n):  # top pole fan
        f.append([top, last + k, last + (k + 1) % n])
    return TriMesh(np.array(v), f)


def _merge_vertices(tris: np.ndarray) -> TriMesh:
    """Indexed mesh from a triangle soup [m, 3, 3] (identical vertices merged)."""
    flat = tris.reshape(-1, 3)
    uniq, inv = np.unique(flat, axis=0, return_inverse=True)
    return TriMesh(uniq, inv.reshape(-1, 3))


def load_stl(path: str) -> TriMesh:
    """Binary or ASCII STL."""
    with open(path, "rb") as fh:
        data = fh.read()
    if len(data) >= 84:
        n = struct.unpack("<I", data[80:84])[0]
        if 84 + 50 * n == len(data):
            rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]),
                                count=n, offset=84)
            return _merge_vertices(rec["v"].astype(np.float64))
    words = data.decode("ascii", errors="replace").split()
    vs = [[float(words[i + 1]), float(words[i + 2]), float(words[i + 3])]
          for i, w in enumerate(words) if w == "vertex"]
    if not vs or len(vs) % 3:
        raise ValueError(f"{path}: not an STL file")
    return _merge_vertices(np.array(vs, np.float64).reshape(-1, 3, 3))


def load_obj(path: str) -> TriMesh:
    """Wavefront OBJ: v records and f records (polygons fan-triangulated,
    1-based or negative indices, texture/normal indices ignored)."""
    vs, fs = [], []
    with open(path) as fh:
        for line in fh:
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                vs.append([float(x) for x in p[1:4]])
            elif p[0] == "f":
                idx = [int(t.split("/")[0]) for t in p[1:]]
                idx = [i - 1 if i > 0 else len(vs) + i for i in idx]
                fs += [[idx[0], idx[k], idx[k + 1]] for k in range(1, len(idx) - 1)]
    return TriMesh(np.array(vs, np.float64), np.array(fs, np.int64))


def load_mesh(path: str) -> TriMesh:
    ext = os.path.splitext(path)[1].lower()
    if ext == ".stl":
        return load_stl(path)
    if ext == ".obj":
        return load_obj(path)
    raise ValueError(f"{path}: unsupported mesh format (STL or OBJ)")


# -------------------------------------------------------------- rotations

def quat_to_matrix(q: Sequence[float]) -> np.ndarray:
    """Rotation matrix of quaternion (w, x, y, z), normalised first."""
    w, x, y, z = np.asarray(q, np.float64) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def matrix_to_quat(R: np.ndarray) -> np.ndarray:
    """Quaternion (w, x, y, z), w >= 0, of a rotation matrix."""
    R = np.asarray(R, np.float64)
    t = np.trace(R)
    if t > 0:
        s = 2.0 * np.sqrt(1.0 + t)
        q = [0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s]
    else:
        i = int(np.argmax(np.diag(R)))
        j, k = (i + 1) % 3, (i + 2) % 3
        s = 2.0 * np.sqrt(1.0 + R[i, i] - R[j, j] - R[k, k])
        q = [0.0] * 4
        q[0] = (R[k, j] - R[j, k]) / s
        q[1 + i] = 0.25 * s
        q[1 + j] = (R[j, i] + R[i, j]) / s
        q[1 + k] = (R[k, i] + R[i, k]) / s
    q = np.array(q)
    return q if q[0] >= 0 else -q


def euler_xyz_deg_to_quat(euler: Sequence[float]) -> np.ndarray:
    """Extrinsic x, then y, then z rotations by the given degrees."""
    ax, ay, az = np.radians(np.asarray(euler, np.float64))
    cx, sx, cy, sy, cz, sz = np.cos(ax), np.sin(ax), np.cos(ay), np.sin(ay), np.cos(az), np.sin(az)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return matrix_to_quat(Rz @ Ry @ Rx)


def _align_z(direction: np.ndarray) -> np.ndarray:
    """Rotation taking +z onto the unit vector `direction` (the smallest one;
    -z maps by a half turn about x)."""
    z = np.array([0.0, 0.0, 1.0])
    c = float(np.clip(direction @ z, -1.0, 1.0))
    if np.allclose(direction, z):
        return np.eye(3)
    if np.allclose(direction, -z):
        return np.diag([1.0, -1.0, -1.0])
    axis = np.cross(z, direction)
    axis /= np.linalg.norm(axis)
    ang = np.arccos(c)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


# ------------------------------------------------------------------ MJCF

def _floats(s: Optional[str], default: str) -> np.ndarray:
    return np.array([float(x) for x in (s if s is not None else default).split()], np.float64)


def _orientation(elem) -> np.ndarray:
    """quat (w, x, y, z) of an element's quat or euler attribute."""
    quat = _floats(elem.get("quat"), "1 0 0 0")
    if elem.get("euler") is not None:
        quat = euler_xyz_deg_to_quat(_floats(elem.get("euler"), "0 0 0"))
    return quat


def _resolve_mesh_path(uri: Optional[str], search_paths: List[str]) -> Optional[str]:
    """mjcf_parser.py:323-338: absolute, else the first search path holding it."""
    if uri is None:
        return None
    if os.path.isabs(uri):
        return uri if os.path.exists(uri) else None
    for sp in search_paths:
        p = os.path.join(sp, uri)
        if os.path.exists(p):
            return p
    return None


def _fromto(sizes, fromto, make) -> TriMesh:
    p1, p2 = fromto[:3], fromto[3:]
    length = float(np.linalg.norm(p2 - p1))
    m = make(sizes[0], length)
    d = (p2 - p1) / length if length > 1e-6 else np.array([0.0, 0.0, 1.0])
    return m.transformed(_align_z(d), 0.5 * (p1 + p2))


def parse_geom(geom, mesh_assets: Dict, search_paths: List[str]) -> Optional[TriMesh]:
    """One geom element as a mesh in its body's frame (mjcf_parser.py:155-299);
    None for an unsupported type or a mesh file that cannot be found."""
    gtype = geom.get("type", "sphere")
    pos = _floats(geom.get("pos"), "0 0 0")
    quat = _orientation(geom)
    fromto = geom.get("fromto")
    m: Optional[TriMesh] = None
    if gtype == "box":
        m = box(2.0 * _floats(geom.get("size"), "0.5 0.5 0.5"))
    elif gtype == "sphere":
        m = icosphere(3, float(_floats(geom.get("size"), "0.5")[0]))
    elif gtype == "ellipsoid":
        m = icosphere(3, 1.0)
        m.vertices *= _floats(geom.get("size"), "0.5 0.5 0.5")
    elif gtype in ("cylinder", "capsule"):
        sizes = _floats(geom.get("size"), "0.5 0.5" if gtype == "cylinder" else "0.5")
        make = (lambda r, h: cylinder(r, h, 32)) if gtype == "cylinder" else (lambda r, h: capsule(h, r, (16, 8)))
        if fromto is not None:
            return _fromto(sizes, _floats(fromto, ""), make)
        half = sizes[1] if len(sizes) > 1 else sizes[0]
        m = make(sizes[0], 2.0 * half)
    elif gtype == "mesh":
        name = geom.get("mesh")
        if name and name in mesh_assets:
            a = mesh_assets[name]
            path = _resolve_mesh_path(a["file"], search_paths)
            if path:
                m = load_mesh(path)
                if not np.allclose(a["scale"], 1.0):
                    m.vertices = m.vertices * a["scale"]
    if m is None:
        return None
    return m.transformed(quat_to_matrix(quat), pos)


def parse_joint(joint, body_name: str) -> Dict:
    """mjcf_parser.py:302-320."""
    rng = _floats(joint.get("range"), "-180 180")
    return {"name": joint.get("name", f"joint_{body_name}"), "type": joint.get("type", "hinge"),
            "body": body_name, "axis": _floats(joint.get("axis"), "0 0 1"),
            "range": (float(rng[0]), float(rng[1])) if len(rng) >= 2 else (-180.0, 180.0)}


def _parse_bodies(parent, parent_name, bodies, joints, assets, search_paths):
    for b in parent.findall("body"):
        name = b.get("name", f"body_{len(bodies)}")
        meshes = [m for m in (parse_geom(g, assets, search_paths) for g in b.findall("geom")) if m is not None]
        if meshes:
            bodies.append({"name": name, "meshes": meshes,
                           "combined_mesh": meshes[0] if len(meshes) == 1 else concatenate(meshes),
                           "parent": parent_name, "pos": _floats(b.get("pos"), "0 0 0"), "quat": _orientation(b)})
        for j in b.findall("joint"):
            joints.append(parse_joint(j, name))
        _parse_bodies(b, name, bodies, joints, assets, search_paths)


def parse_mjcf(mjcf_path: str, mesh_search_paths: Optional[List[str]] = None) -> Dict:
    """Bodies (with geometry) and joints of an MJCF file (mjcf_parser.py:15-95):
    {'model_name', 'bodies': [{'name', 'meshes', 'combined_mesh', 'parent',
    'pos', 'quat'}], 'joints': [{'name', 'type', 'body', 'axis', 'range'}]};
    pos/quat are relative to the parent body. Bodies without geometry are not
    listed but their children are (their parent is the geometry-less body)."""
    if not os.path.exists(mjcf_path):
        raise FileNotFoundError(f"MJCF file not found: {mjcf_path}")
    search = list(mesh_search_paths or [])
    mjcf_dir = os.path.dirname(os.path.abspath(mjcf_path))
    if mjcf_dir not in search:
        search.insert(0, mjcf_dir)
    root = ET.parse(mjcf_path).getroot()
    comp = root.find("compiler")
    if comp is not None and comp.get("meshdir") is not None:
        md = os.path.normpath(os.path.join(mjcf_dir, comp.get("meshdir")))
        if md not in search:
            search.insert(0, md)
    assets = {}
    asset = root.find("asset")
    if asset is not None:
        for m in asset.findall("mesh"):
            assets[m.get("name")] = {"file": m.get("file"), "scale": _floats(m.get("scale"), "1 1 1")}
    bodies: List[dict] = []
    joints: List[dict] = []
    wb = root.find("worldbody")
    if wb is not None:
        _parse_bodies(wb, None, bodies, joints, assets, search)
    return {"model_name": root.get("model", "unnamed_model"), "bodies": bodies, "joints": joints}


def get_body_world_positions(mjcf_data: Dict) -> Dict[str, np.ndarray]:
    """World position of every listed body (mjcf_parser.py:341-372): its
    position carried up the chain of listed parents."""
    return {k: v[0] for k, v in get_body_world_poses(mjcf_data).items()}


def get_body_world_poses(mjcf_data: Dict) -> Dict[str, tuple]:
    """(world position, world rotation matrix) of every listed body, composed
    through its chain of listed parents."""
    bmap = {b["name"]: b for b in mjcf_data["bodies"]}
    out = {}
    for b in mjcf_data["bodies"]:
        p, R = b["pos"].copy(), quat_to_matrix(b["quat"])
        parent = b["parent"]
        while parent is not None and parent in bmap:
            pb = bmap[parent]
            Rp = quat_to_matrix(pb["quat"])
            p = Rp @ p + pb["pos"]
            R = Rp @ R
            parent = pb["parent"]
        out[b["name"]] = (p, R)
    return out


def parse_mjcf_to_meshes(mjcf_path: str, mesh_search_paths: Optional[List[str]] = None) -> Dict[str, TriMesh]:
    """{body name: combined mesh} (mjcf_parser.py:375-393)."""
    d = parse_mjcf(mjcf_path, mesh_search_paths)
    return {b["name"]: b["combined_mesh"] for b in d["bodies"] if b["combined_mesh"] is not None}


def parse_mjcf_as_urdf_format(mjcf_path: str) -> Dict:
    """The parse_urdf-shaped view (mjcf_parser.py:396-459): one link per body
    (mass 1, inertia 0.01 I, geometry offsets kept in the mesh) and one
    revolute joint about z per parent-child pair, its origin the child's
    parent-relative pose (rpy: extrinsic x-y-z radians)."""
    d = parse_mjcf(mjcf_path)
    links = [{"name": b["name"], "mesh": b["combined_mesh"], "origin_xyz": np.zeros(3), "origin_rpy": np.zeros(3),
              "mass": 1.0, "inertia": np.eye(3) * 0.01} for b in d["bodies"]]
    joints = []
    for b in d["bodies"]:
        if b["parent"] is not None:
            R = quat_to_matrix(b["quat"])
            rpy = np.array([np.arctan2(R[2, 1], R[2, 2]), np.arcsin(-np.clip(R[2, 0], -1.0, 1.0)),
                            np.arctan2(R[1, 0], R[0, 0])])
            joints.append({"name": f"joint_{b['parent']}_{b['name']}", "type": "revolute", "parent": b["parent"],
                           "child": b["name"], "origin_xyz": b["pos"], "origin_rpy": rpy,
                           "axis": np.array([0, 0, 1])})
    return {"robot_name": d["model_name"], "links": links, "joints": joints}


# ------------------------------------------- kinematic model (no MuJoCo)

@dataclass
class Geom:
    """A geom in the model's reference configuration: type, MuJoCo size
    (fromto capsules and cylinders: radius and half length), world position
    and rotation (fromto geoms: z along the segment, the smallest rotation
    from +z), its body index, density and bounding radius."""

    name: str
    body: int
    type: str
    size: np.ndarray
    pos: np.ndarray
    mat: np.ndarray
    density: float
    rbound: float
    mesh: Optional[TriMesh] = None  # mesh geoms: vertices in the geom frame


@dataclass
class Body:
    name: str
    parent: int          # -1: the world
    pos: np.ndarray      # world position of the body frame
    mat: np.ndarray      # world rotation of the body frame
    quat: np.ndarray     # the same rotation as (w, x, y, z)
    ipos: np.ndarray     # world position of the centre of mass


@dataclass
class Model:
    """Bodies and geoms of an MJCF file at its reference configuration
    (every joint at 0: the pose mj_forward gives for qpos0), with the centre
    of mass of each body from its geoms (uniform density per geom, volumes of
    the exact shapes, as MuJoCo's inertiafromgeom) or its inertial element."""

    name: str
    bodies: List[Body]
    geoms: List[Geom]
    joints: List[dict]

    def body_id(self, name: str) -> int:
        for i, b in enumerate(self.bodies):
            if b.name == name:
                return i
        raise KeyError(f"no body named {name!r}")

    def geom_id(self, name: str) -> int:
        for i, g in enumerate(self.geoms):
            if g.name == name:
                return i
        raise KeyError(f"no geom named {name!r}")


def _geom_volume(g: Geom) -> float:
    r = g.size[0]
    if g.type == "sphere":
        return 4.0 / 3.0 * np.pi * r ** 3
    if g.type == "capsule":
        return np.pi * r * r * 2.0 * g.size[1] + 4.0 / 3.0 * np.pi * r ** 3
    if g.type == "cylinder":
        return np.pi * r * r * 2.0 * g.size[1]
    if g.type == "box":
        return 8.0 * g.size[0] * g.size[1] * g.size[2]
    if g.type == "ellipsoid":
        return 4.0 / 3.0 * np.pi * g.size[0] * g.size[1] * g.size[2]
    if g.type == "mesh" and g.mesh is not None:
        return abs(g.mesh.volume)
    return 0.0


def load_model(mjcf_path: str, mesh_search_paths: Optional[List[str]] = None) -> Model:
    """The kinematic model of an MJCF file at its reference configuration."""
    if not os.path.exists(mjcf_path):
        raise FileNotFoundError(f"MJCF file not found: {mjcf_path}")
    search = list(mesh_search_paths or [])
    mjcf_dir = os.path.dirname(os.path.abspath(mjcf_path))
    if mjcf_dir not in search:
        search.insert(0, mjcf_dir)
    root = ET.parse(mjcf_path).getroot()
    comp = root.find("compiler")
    if comp is not None and comp.get("meshdir") is not None:
        search.insert(0, os.path.normpath(os.path.join(mjcf_dir, comp.get("meshdir"))))
    assets = {}
    asset = root.find("asset")
    if asset is not None:
        for m in asset.findall("mesh"):
            assets[m.get("name")] = {"file": m.get("file"), "scale": _floats(m.get("scale"), "1 1 1")}
    bodies: List[Body] = []
    geoms: List[Geom] = []
    joints: List[dict] = []

    def add_geoms(elem, bid, P, R):
        for g in elem.findall("geom"):
            gtype = g.get("type", "sphere")
            if gtype == "plane":
                continue
            size = _floats(g.get("size"), "0.5 0.5 0.5")
            fromto = g.get("fromto")
            mesh = None
            if fromto is not None and gtype in ("capsule", "cylinder"):
                ft = _floats(fromto, "")
                p1, p2 = ft[:3], ft[3:]
                length = float(np.linalg.norm(p2 - p1))
                d = (p2 - p1) / length if length > 1e-6 else np.array([0.0, 0.0, 1.0])
                lpos, lmat = 0.5 * (p1 + p2), _align_z(d)
                size = np.array([size[0], 0.5 * length, 0.0])
            else:
                lpos, lmat = _floats(g.get("pos"), "0 0 0"), quat_to_matrix(_orientation(g))
                size = np.concatenate([size, np.zeros(3)])[:3]
                if gtype in ("capsule", "cylinder") and len(_floats(g.get("size"), "0.5")) == 1:
                    size[1] = size[0]
                if gtype == "mesh":
                    name = g.get("mesh")
                    if name in assets:
                        path = _resolve_mesh_path(assets[name]["file"], search)
                        if path:
                            mesh = load_mesh(path)
                            mesh.vertices = mesh.vertices * assets[name]["scale"]
            r = size[0]
            rb = {"sphere": r, "capsule": r + size[1], "cylinder": float(np.hypot(r, size[1])),
                  "box": float(np.linalg.norm(size)), "ellipsoid": float(size.max())}.get(gtype, 0.0)
            if mesh is not None:
                rb = float(np.linalg.norm(mesh.vertices, axis=1).max())
            geoms.append(Geom(g.get("name", f"geom_{len(geoms)}"), bid, gtype, size, P + R @ lpos, R @ lmat,
                              float(g.get("density", "1000")), rb, mesh))

    def walk(elem, parent, P, R):
        for b in elem.findall("body"):
            lp = _floats(b.get("pos"), "0 0 0")
            Pb, Rb = P + R @ lp, R @ quat_to_matrix(_orientation(b))
            bid = len(bodies)
            bodies.append(Body(b.get("name", f"body_{bid}"), parent, Pb, Rb, matrix_to_quat(Rb), Pb.copy()))
            first = len(geoms)
            add_geoms(b, bid, Pb, Rb)
            inertial = b.find("inertial")
            if inertial is not None:
                bodies[bid].ipos = Pb + Rb @ _floats(inertial.get("pos"), "0 0 0")
            elif len(geoms) > first:
                w = np.array([g.density * _geom_volume(g) for g in geoms[first:]])
                c = np.array([g.pos for g in geoms[first:]])
                if w.sum() > 0:
                    bodies[bid].ipos = (w[:, None] * c).sum(0) / w.sum()
            for j in b.findall("joint"):
                joints.append(parse_joint(j, bodies[bid].name))
            walk(b, bid, Pb, Rb)

    wb = root.find("worldbody")
    if wb is not None:
        walk(wb, -1, np.zeros(3), np.eye(3))
    return Model(root.get("model", "unnamed_model"), bodies, geoms, joints)


# ------------------------------------------------- planar outlines (2D)

def resample_polygon(vertices: np.ndarray, n_samples: int) -> np.ndarray:
    """n_samples points evenly spaced along the closed boundary of the convex
    hull of `vertices` (envs/lbm/mujoco_to_lbm.py:46-97)."""
    from scipy.spatial import ConvexHull

    v = np.asarray(vertices, np.float64).reshape(-1, 2)
    if len(v) < 2:
        return v
    if len(v) >= 3:
        try:
            v = v[ConvexHull(v).vertices]
        except Exception:  # degenerate (collinear) input keeps its points
            pass
    closed = np.vstack([v, v[:1]])
    seg = np.diff(closed, axis=0)
    seg_len = np.linalg.norm(seg, axis=1)
    cum = np.concatenate([[0.0], np.cumsum(seg_len)])
    total = cum[-1]
    if total < 1e-10:
        a = np.linspace(0.0, 2.0 * np.pi, n_samples, endpoint=False)
        return np.column_stack([0.1 * np.cos(a), 0.1 * np.sin(a)])
    out = []
    for s in np.linspace(0.0, total, n_samples, endpoint=False):
        k = min(int(np.searchsorted(cum[1:], s)), len(v) - 1)
        t = float(np.clip((s - cum[k]) / (seg_len[k] + 1e-10), 0.0, 1.0))
        out.append(v[k] + t * seg[k])
    return np.array(out)


def geom_outline_2d(model: Model, geom: int, n_samples: int = 20) -> np.ndarray:
    """Outline of one geom projected onto the x-y plane in world coordinates
    (get_geom_vertices_2d, envs/lbm/mujoco_to_lbm.py:150-271): box corners;
    circles of n_samples points for spheres and cylinders; a capsule as two
    half circles of n_samples // 2 points around its axis's projection; the
    x-y ellipse of an ellipsoid; a mesh's projected vertices (resampled to
    n_samples along their hull); then turned by the geom's rotation about z
    (from its x axis) and moved to its position. As in the reference, the
    capsule's half circles are laid along local +y turned by the angle of its
    projected axis d measured from +y, that is along (-d_x, d_y), before that
    rotation: capsules along the x or y axis (the reference's eels) come out
    along their axis, oblique ones do not."""
    g = model.geoms[geom]
    ang = np.arctan2(g.mat[1, 0], g.mat[0, 0])
    circle = np.linspace(0.0, 2.0 * np.pi, n_samples, endpoint=False)
    if g.type == "box":
        hx, hy = g.size[0], g.size[1]
        local = np.array([[-hx, -hy], [hx, -hy], [hx, hy], [-hx, hy]])
    elif g.type in ("sphere", "cylinder"):
        local = np.column_stack([g.size[0] * np.cos(circle), g.size[0] * np.sin(circle)])
    elif g.type == "capsule":
        r, half = g.size[0], g.size[1]
        if half < 1e-6:
            half = max(g.rbound - r, 0.0)
        axis = g.mat[:2, 2]
        n = np.linalg.norm(axis)
        cap_ang = np.arctan2(axis[0] / n, axis[1] / n) if n > 1e-6 else 0.0
        nh = n_samples // 2
        top = np.linspace(0.0, np.pi, nh, endpoint=False)
        bot = np.linspace(np.pi, 2.0 * np.pi, nh, endpoint=False)
        cap = np.vstack([np.column_stack([r * np.cos(top), half + r * np.sin(top)]),
                         np.column_stack([r * np.cos(bot), -half + r * np.sin(bot)])])
        c, s = np.cos(cap_ang), np.sin(cap_ang)
        local = cap @ np.array([[c, -s], [s, c]]).T
    elif g.type == "ellipsoid":
        local = np.column_stack([g.size[0] * np.cos(circle), g.size[1] * np.sin(circle)])
    elif g.type == "mesh" and g.mesh is not None:
        local = g.mesh.vertices[:, :2]
        if n_samples is not None and len(local) != n_samples:
            local = resample_polygon(local, n_samples)
    else:  # unsupported: a small circle, as the reference does
        local = np.column_stack([0.1 * np.cos(circle), 0.1 * np.sin(circle)])
    c, s = np.cos(ang), np.sin(ang)
    return local @ np.array([[c, -s], [s, c]]).T + g.pos[:2]


def body_outline_2d(model: Model, body: str, n_samples: int = 20) -> np.ndarray:
    """Every geom outline of a body, stacked (get_body_vertices_2d,
    mujoco_to_lbm.py:274-307)."""
    bid = model.body_id(body)
    parts = [geom_outline_2d(model, k, n_samples) for k, g in enumerate(model.geoms) if g.body == bid]
    if not parts:
        raise ValueError(f"body {body!r} has no geometry")
    return np.vstack(parts)


def polygon_from_model(model: Model, name: str, n_samples: int = 20, use_convex_hull: bool = True,
                       normalize: bool = True, is_body: bool = True) -> Dict:
    """{'vertices', 'position', 'angle'} of a body or geom for the 2D solver
    (extract_lbm_polygon_from_mujoco, mujoco_to_lbm.py:343-395): the outline,
    its convex hull (counter-clockwise), centred on its vertex mean; the
    body's centre of mass (or the geom's position) and rotation about z."""
    from scipy.spatial import ConvexHull

    if is_body:
        v = body_outline_2d(model, name, n_samples)
        b = model.bodies[model.body_id(name)]
        pos = b.ipos[:2].copy()
        w, x, y, z = b.quat
        ang = float(np.arctan2(2.0 * (w * z + x * y), 1.0 - 2.0 * (y * y + z * z)))
    else:
        gid = model.geom_id(name)
        v = geom_outline_2d(model, gid, n_samples)
        g = model.geoms[gid]
        pos = g.pos[:2].copy()
        ang = float(np.arctan2(g.mat[1, 0], g.mat[0, 0]))
    if use_convex_hull and len(v) > 3:
        try:
            v = v[ConvexHull(v).vertices]
        except Exception:  # degenerate input keeps its points
            pass
    if normalize:
        v = v - v.mean(axis=0)
    return {"vertices": v, "position": pos, "angle": ang}

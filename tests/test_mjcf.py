"""MJCF geometry (openhomelbm_amd.mjcf: the reference's mjcf_parser.py and
the planar outlines of mujoco_to_lbm.py, without trimesh or MuJoCo), pinned
by closed forms:

* meshes: box volumes exactly; every icosphere, cylinder and capsule vertex
  on the exact surface; the prism's volume n/2 r^2 sin(2 pi/n) h exactly;
  inscribed volumes between the exact body and its inner parallel body;
  closed and outward-oriented (every edge met once each way);
* MJCF: body tree, parent-relative poses, euler (extrinsic x-y-z degrees)
  against hand-built quaternions, fromto capsules on their segment, world
  positions of a rotated three-body chain by hand, STL (binary and ASCII)
  and OBJ assets, the URDF-shaped view;
* voxel counts: a box body whose faces lie between cell centres fills
  exactly its cells through the mesh voxelizer oracle;
* 2D outlines: a capsule's hull extents, centre of mass and angle; a rotated
  box's corners; circles and ellipses; evenly resampled hulls.
"""
import os
import struct

import numpy as np
import pytest

from openhomelbm_amd import mjcf
from oracle import mesh as mesh_oracle


def _seg_dist(P, a, b):
    ab = b - a
    L2 = ab @ ab
    s = np.zeros(len(P)) if L2 == 0 else np.clip((P - a) @ ab / L2, 0.0, 1.0)
    return np.linalg.norm(P - (a + s[:, None] * ab), axis=1)


def _capsule_volume(r, h):
    return np.pi * r * r * h + 4.0 / 3.0 * np.pi * r ** 3


# ----------------------------------------------------------------- meshes

def test_box_mesh():
    m = mjcf.box((2.0, 3.0, 5.0))
    assert m.is_watertight and m.volume == pytest.approx(30.0, abs=1e-12)
    assert np.allclose(np.abs(m.vertices), [1.0, 1.5, 2.5])


def test_icosphere_mesh():
    m = mjcf.icosphere(3, 2.5)
    assert len(m.vertices) == 642 and len(m.faces) == 1280 and m.is_watertight
    assert np.allclose(np.linalg.norm(m.vertices, axis=1), 2.5, atol=1e-12)
    exact = 4.0 / 3.0 * np.pi * 2.5 ** 3
    assert 0.99 * exact < m.volume < exact


def test_cylinder_mesh():
    r, h, n = 1.7, 3.1, 32
    m = mjcf.cylinder(r, h, n)
    assert m.is_watertight
    assert m.volume == pytest.approx(0.5 * n * r * r * np.sin(2 * np.pi / n) * h, rel=1e-12)
    side = np.abs(np.abs(m.vertices[:, 2]) - h / 2) < 1e-12
    ring = np.linalg.norm(m.vertices[:, :2], axis=1)
    assert side.all() and np.allclose(ring[ring > 1e-9], r, atol=1e-12)


def test_capsule_mesh():
    r, h = 1.3, 2.2
    m = mjcf.capsule(h, r, (16, 8))
    assert m.is_watertight
    d = _seg_dist(m.vertices, np.array([0, 0, -h / 2]), np.array([0, 0, h / 2]))
    assert np.allclose(d, r, atol=1e-12)
    delta = r * (1 - np.cos(np.pi / 8))  # the largest sagitta of a 22.5-degree patch
    assert _capsule_volume(r - delta, h) < m.volume < _capsule_volume(r, h)


# ------------------------------------------------------------------- MJCF

def _write_stl_binary(path, tris):
    with open(path, "wb") as f:
        f.write(b"\0" * 80 + struct.pack("<I", len(tris)))
        for t in tris:
            f.write(struct.pack("<12fH", 0, 0, 0, *np.asarray(t, np.float32).ravel(), 0))


def _write_stl_ascii(path, tris):
    with open(path, "w") as f:
        f.write("solid s\n")
        for t in tris:
            f.write("facet normal 0 0 0\n outer loop\n")
            for v in t:
                f.write(f"  vertex {v[0]} {v[1]} {v[2]}\n")
            f.write(" endloop\nendfacet\n")
        f.write("endsolid s\n")


SCENE = """<mujoco model="probe">
  <compiler angle="degree" meshdir="assets"/>
  <asset>
    <mesh name="cube_b" file="cube_b.stl" scale="2 2 2"/>
    <mesh name="cube_a" file="cube_a.stl"/>
    <mesh name="cube_o" file="cube.obj"/>
  </asset>
  <worldbody>
    <geom type="plane" size="5 5 0.1"/>
    <body name="base" pos="1 2 3" euler="0 0 90">
      <joint name="free" type="free"/>
      <geom type="box" size="0.5 1 1.5"/>
      <body name="arm" pos="2 0 0">
        <joint name="hinge1" type="hinge" axis="0 0 1" range="-45 45"/>
        <geom type="capsule" fromto="0 0 0  0 -3 0" size="0.25"/>
        <body name="hand" pos="0 -3 0" quat="0.7071067811865476 0.7071067811865476 0 0">
          <joint name="hinge2" axis="0 1 0"/>
          <geom type="sphere" size="0.4" pos="0 0 1"/>
          <geom type="cylinder" size="0.2 0.6" euler="90 0 0"/>
        </body>
      </body>
    </body>
    <body name="meshes" pos="-4 0 0">
      <geom type="mesh" mesh="cube_b"/>
      <geom type="mesh" mesh="cube_a" pos="0 0 3"/>
      <geom type="mesh" mesh="cube_o" pos="0 0 -3"/>
      <geom type="ellipsoid" size="1 2 3" pos="5 0 0"/>
    </body>
    <body name="empty" pos="0 9 0">
      <body name="under_empty" pos="1 0 0"><geom type="sphere" size="1"/></body>
    </body>
  </worldbody>
</mujoco>
"""


@pytest.fixture
def scene(tmp_path):
    os.makedirs(tmp_path / "assets")
    cube = mjcf.box((1.0, 1.0, 1.0))
    tris = cube.vertices[cube.faces]
    _write_stl_binary(tmp_path / "assets" / "cube_b.stl", tris)
    _write_stl_ascii(tmp_path / "assets" / "cube_a.stl", tris)
    with open(tmp_path / "assets" / "cube.obj", "w") as f:
        for v in cube.vertices:
            f.write(f"v {v[0]} {v[1]} {v[2]}\n")
        for a, b, c in cube.faces + 1:
            f.write(f"f {a}/1 {b}/1 {c}/1\n")
    p = tmp_path / "probe.xml"
    p.write_text(SCENE)
    return str(p)


def test_parse_tree_and_poses(scene):
    d = mjcf.parse_mjcf(scene)
    assert d["model_name"] == "probe"
    names = [b["name"] for b in d["bodies"]]
    assert names == ["base", "arm", "hand", "meshes", "under_empty"]  # "empty" has no geometry
    parents = {b["name"]: b["parent"] for b in d["bodies"]}
    assert parents == {"base": None, "arm": "base", "hand": "arm", "meshes": None, "under_empty": "empty"}
    base = d["bodies"][0]
    c = np.cos(np.pi / 4)
    assert np.allclose(base["quat"], [c, 0, 0, c])  # euler z 90 degrees
    assert [j["name"] for j in d["joints"]] == ["free", "hinge1", "hinge2"]
    j1 = d["joints"][1]
    assert j1["type"] == "hinge" and j1["range"] == (-45.0, 45.0) and j1["body"] == "arm"
    assert d["joints"][2]["type"] == "hinge" and d["joints"][2]["range"] == (-180.0, 180.0)
    # euler angles are extrinsic x, y, z
    q = mjcf.euler_xyz_deg_to_quat([90, 0, 90])
    R = mjcf.quat_to_matrix(q)
    Rx = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]])
    Rz = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]])
    assert np.allclose(R, Rz @ Rx)


def test_geoms(scene):
    d = mjcf.parse_mjcf(scene)
    b = {x["name"]: x for x in d["bodies"]}
    assert b["base"]["combined_mesh"].volume == pytest.approx(1 * 2 * 3, abs=1e-12)
    arm = b["arm"]["combined_mesh"]
    assert np.allclose(_seg_dist(arm.vertices, np.zeros(3), np.array([0, -3.0, 0])), 0.25, atol=1e-12)
    assert arm.is_watertight and 0 < arm.volume < _capsule_volume(0.25, 3.0)
    hand = b["hand"]
    assert len(hand["meshes"]) == 2 and hand["combined_mesh"].faces.shape[0] == 1280 + 128
    sph, cyl = hand["meshes"]
    assert np.allclose(np.linalg.norm(sph.vertices - [0, 0, 1], axis=1), 0.4, atol=1e-12)
    # euler x 90: the cylinder's axis (z) turns onto -y
    assert np.allclose(np.abs(cyl.vertices[:, 1]).max(), 0.6) and np.allclose(np.hypot(cyl.vertices[:, 0], cyl.vertices[:, 2]).max(), 0.2)
    ms = b["meshes"]["meshes"]
    assert len(ms) == 4
    assert ms[0].volume == pytest.approx(8.0, abs=1e-9)  # binary STL, scale 2
    assert ms[1].volume == pytest.approx(1.0, abs=1e-9) and np.allclose(ms[1].vertices.mean(0), [0, 0, 3])
    assert ms[2].volume == pytest.approx(1.0, abs=1e-12) and np.allclose(ms[2].vertices.mean(0), [0, 0, -3])
    assert all(m.is_watertight for m in ms[:3])
    e = ms[3].vertices - [5, 0, 0]
    assert np.allclose((e[:, 0] / 1) ** 2 + (e[:, 1] / 2) ** 2 + (e[:, 2] / 3) ** 2, 1.0, atol=1e-12)


def test_world_positions_and_urdf_view(scene):
    d = mjcf.parse_mjcf(scene)
    wp = mjcf.get_body_world_positions(d)
    Rz = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]])
    assert np.allclose(wp["base"], [1, 2, 3])
    arm = np.array([1, 2, 3]) + Rz @ np.array([2, 0, 0])
    assert np.allclose(wp["arm"], arm)
    assert np.allclose(wp["hand"], arm + Rz @ np.array([0, -3, 0]))
    assert np.allclose(wp["under_empty"], [1, 0, 0])  # the geometry-less parent is not listed
    poses = mjcf.get_body_world_poses(d)
    Rx = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]])
    assert np.allclose(poses["hand"][1], Rz @ Rx)
    u = mjcf.parse_mjcf_as_urdf_format(scene)
    assert u["robot_name"] == "probe" and len(u["links"]) == 5
    j = {x["child"]: x for x in u["joints"]}
    assert set(j) == {"arm", "hand", "under_empty"}
    assert np.allclose(j["hand"]["origin_rpy"], [np.pi / 2, 0, 0])
    assert mjcf.parse_mjcf_to_meshes(scene).keys() == {"base", "arm", "hand", "meshes", "under_empty"}


def test_box_body_voxel_count(tmp_path):
    """A box body whose faces lie between cell centres fills exactly its
    cells (the mesh voxelizer oracle on the posed triangles)."""
    p = tmp_path / "box.xml"
    p.write_text('<mujoco><worldbody><body name="b" pos="10 9 8">'
                 '<geom type="box" size="2.5 1.5 3.5"/></body></worldbody></mujoco>')
    d = mjcf.parse_mjcf(str(p))
    (pos, R), = mjcf.get_body_world_poses(d).values()
    m = d["bodies"][0]["combined_mesh"].transformed(R, pos)
    mask = mesh_oracle.voxelize((20, 20, 20), m.vertices[m.faces].reshape(-1, 9), np.zeros(len(m.faces), np.int64))
    assert mask.sum() == 5 * 3 * 7
    z, y, x = np.nonzero(mask)
    assert (x.min(), x.max(), y.min(), y.max(), z.min(), z.max()) == (8, 12, 8, 10, 5, 11)


def test_missing_file_and_unknown_geom(tmp_path):
    with pytest.raises(FileNotFoundError):
        mjcf.parse_mjcf(str(tmp_path / "nope.xml"))
    p = tmp_path / "m.xml"
    p.write_text('<mujoco><worldbody><body name="b"><geom type="hfield"/>'
                 '<geom type="mesh" mesh="absent"/></body></worldbody></mujoco>')
    assert mjcf.parse_mjcf(str(p))["bodies"] == []


# ------------------------------------------------------- planar outlines

EEL = """<mujoco model="chain">
  <worldbody>
    <body name="s1" pos="0 0 0.05">
      <joint type="free"/>
      <geom name="g1" type="capsule" fromto="0 0 0  0 -0.1 0" size="0.03" density="400"/>
      <body name="s2" pos="0 -0.1 0">
        <joint type="hinge" axis="0 0 1"/>
        <geom name="g2" type="capsule" fromto="0 0 0  0.08 -0.06 0" size="0.02"/>
        <geom name="g2b" type="sphere" size="0.02" pos="0.08 -0.06 0" density="4000"/>
      </body>
    </body>
    <body name="plate" pos="1 1 0" euler="0 0 30">
      <geom name="pb" type="box" size="0.2 0.1 0.05"/>
    </body>
    <body name="disc" pos="-1 0 0"><geom name="pc" type="cylinder" size="0.3 0.1"/>
      <geom name="pe" type="ellipsoid" size="0.5 0.2 0.1" pos="2 0 0"/></body>
  </worldbody>
</mujoco>
"""


def test_capsule_outline(tmp_path):
    p = tmp_path / "eel.xml"
    p.write_text(EEL)
    m = mjcf.load_model(str(p))
    info = mjcf.polygon_from_model(m, "s1", n_samples=20)
    v = info["vertices"]
    # the hull of two half circles of 10 points about y = 0 and y = -0.1
    assert info["angle"] == 0.0 and np.allclose(info["position"], [0.0, -0.05])
    assert np.allclose(v[:, 0].max() - v[:, 0].min(), 2 * 0.03)
    assert np.allclose(v[:, 1].max() - v[:, 1].min(), 0.1 + 2 * 0.03, atol=1e-12)
    assert abs(v.mean(axis=0)).max() < 1e-12
    ccw = np.sum(v[:, 0] * np.roll(v[:, 1], -1) - np.roll(v[:, 0], -1) * v[:, 1]) > 0
    assert ccw
    # the second body: centre of mass of a capsule and a sphere, weighted by density x volume
    b2 = m.bodies[m.body_id("s2")]
    wc = 1000 * _capsule_volume(0.02, 0.1)
    ws = 4000 * 4.0 / 3.0 * np.pi * 0.02 ** 3
    com = (wc * np.array([0.04, -0.13, 0.05]) + ws * np.array([0.08, -0.16, 0.05])) / (wc + ws)
    assert np.allclose(b2.ipos, com)
    # an oblique capsule, axis d = (0.8, -0.6): the reference's formula lays
    # the outline along (-d_x, d_y) turned by the angle of the geom's x axis
    # (here (0.36, 0.48, -0.8), the smallest rotation taking z onto d), that is
    # along -y; centred on the segment's midpoint, half length + radius each way
    g = mjcf.geom_outline_2d(m, m.geom_id("g2"), 20) - [0.04, -0.13]
    assert np.isclose(g[:, 1].max(), 0.07) and np.isclose(g[:, 1].min(), -0.07)
    assert np.isclose(np.abs(g[:, 0]).max(), 0.02)


def test_box_circle_and_ellipse_outlines(tmp_path):
    p = tmp_path / "eel.xml"
    p.write_text(EEL)
    m = mjcf.load_model(str(p))
    info = mjcf.polygon_from_model(m, "pb", is_body=False)
    a = np.radians(30)
    assert np.isclose(info["angle"], a) and np.allclose(info["position"], [1, 1])
    R = np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]])
    corners = np.array([[-0.2, -0.1], [0.2, -0.1], [0.2, 0.1], [-0.2, 0.1]]) @ R.T
    assert len(info["vertices"]) == 4
    assert all(np.isclose(np.linalg.norm(info["vertices"] - c, axis=1), 0).any() for c in corners)
    disc = mjcf.geom_outline_2d(m, m.geom_id("pc"), 24)
    assert np.allclose(np.linalg.norm(disc - [-1, 0], axis=1), 0.3)
    ell = mjcf.geom_outline_2d(m, m.geom_id("pe"), 24) - [1, 0]
    assert np.allclose((ell[:, 0] / 0.5) ** 2 + (ell[:, 1] / 0.2) ** 2, 1.0)


def test_resample_polygon():
    sq = np.array([[0, 0], [2, 0], [2, 2], [0, 2], [1, 1]], float)  # an interior point is dropped by the hull
    r = mjcf.resample_polygon(sq, 8)
    assert len(r) == 8
    d = np.linalg.norm(np.diff(np.vstack([r, r[:1]]), axis=0), axis=1)
    assert np.allclose(d, 1.0)  # perimeter 8, eight even steps along the edges

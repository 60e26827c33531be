"""The C-ABI library loads and exports every declared entry point; argument
errors are reported without touching the GPU (runs on a CPU-only machine)."""
import ctypes as C
import os
import re

import pytest

from openhomelbm_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?[\w\*]+\s+\**(hl_\w+)\s*\(", text, re.M))
    return names


def test_library_exports_every_declared_symbol():
    lib = N.load()
    declared = _declared()
    assert {"hl_step", "hl_forces", "hl_voxelize", "hl_mass", "hl_init_uniform"} <= declared
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert set(N.EXPORTS) <= declared | {"hl_last_error"}


def test_env_layout_matches_binding():
    lib = N.load()
    assert lib.hl_env_bytes() == C.sizeof(N.HLEnv)
    assert lib.hl_version() >= 100


@pytest.mark.parametrize("field,value,msg", [
    ("lattice", 5, "lattice"),
    ("nx", 2, "extent"),
    ("nu", -1.0, "viscosity"),
    ("dtype", 9, "dtype"),
])
def test_invalid_env_is_rejected_before_any_gpu_call(field, value, msg):
    lib = N.load()
    e = N.HLEnv()
    e.lattice, e.dtype, e.nx, e.ny, e.nz = 27, 0, 8, 8, 8
    e.stride = 512
    e.moments[0] = 0x1000
    e.moments[1] = 0x2000
    e.nu = 0.1
    setattr(e, field, value)
    rc = lib.hl_step(C.byref(e), 0, 1, 0, None, None)
    assert rc == -1
    assert msg in lib.hl_last_error().decode()
    with pytest.raises(N.HLError):
        N.check(rc, "hl_step")


def test_bad_bc_type_rejected():
    lib = N.load()
    e = N.HLEnv()
    e.lattice, e.dtype, e.nx, e.ny, e.nz = 9, 0, 8, 8, 1
    e.stride = 64
    e.moments[0] = 0x1000
    e.moments[1] = 0x2000
    e.nu = 0.1
    e.bc_type[2] = 7
    assert lib.hl_step(C.byref(e), 0, 1, 0, None, None) == -1
    assert "bc_type" in lib.hl_last_error().decode()


def test_batch_rejects_unsupported_environments():
    lib = N.load()
    e = N.HLEnv()
    e.lattice, e.dtype, e.nx, e.ny, e.nz = 27, 0, 8, 8, 8
    e.stride = 512
    e.moments[0] = 0x1000
    e.moments[1] = 0x2000
    e.nu = 0.1
    h = C.c_void_p()
    assert lib.hl_batch_create(C.byref(e), 1, C.byref(h)) == -2
    assert "D2Q9 FP32" in lib.hl_last_error().decode()
    assert not h.value
    assert lib.hl_batch_create(C.byref(e), 0, C.byref(h)) == -1
    assert lib.hl_batch_step(None, None, 1, None, 0, None) == -1
    assert lib.hl_batch_replay(None, None) == -1
    assert "NULL" in lib.hl_last_error().decode()


def _d2q9_env(nsolid=1, dtype=0):
    e = N.HLEnv()
    e.lattice, e.dtype, e.nx, e.ny, e.nz = 9, dtype, 8, 8, 1
    e.stride = 64
    e.moments[0] = 0x1000
    e.moments[1] = 0x2000
    e.nu = 0.1
    e.nsolid = nsolid
    e.solids = 0x3000 if nsolid else None
    return e


def test_polygon_worlds_arguments_rejected_before_any_gpu_call():
    """hl_step_polygons_worlds checks every world's descriptor and tables
    and names the world at fault, before any device work."""
    lib = N.load()
    P = C.c_void_p

    def call(envs, n, verts=(0x10, 0x10), counts=(0x20, 0x20), maxv=(4, 4), r2=(0x30, 0x30), cur=0, nsteps=1):
        arr = (N.HLEnv * len(envs))(*envs) if envs is not None else None
        return lib.hl_step_polygons_worlds(arr, n, cur, nsteps, (P * 2)(*verts), None, (P * 2)(*counts),
                                           (C.c_int32 * 2)(*maxv), (P * 2)(*r2), None, None)

    two = [_d2q9_env(), _d2q9_env()]
    assert call(None, 2) == -1 and "envs is NULL" in lib.hl_last_error().decode()
    assert call(two, 0) == -1 and "nworld" in lib.hl_last_error().decode()
    assert call(two, 2, cur=2) == -1 and "cur" in lib.hl_last_error().decode()
    assert call(two, 2, nsteps=-1) == -1 and "nsteps" in lib.hl_last_error().decode()
    d3 = _d2q9_env()
    d3.lattice, d3.nz, d3.stride = 27, 8, 512
    assert call([_d2q9_env(), d3], 2) == -2 and "world 1" in lib.hl_last_error().decode()
    assert call([_d2q9_env(), _d2q9_env(dtype=1)], 2) == -1 and "same dtype" in lib.hl_last_error().decode()
    assert call([_d2q9_env(), _d2q9_env(nsolid=0)], 2) == -1 and "nsolid" in lib.hl_last_error().decode()
    assert call(two, 2, maxv=(4, 0)) == -1 and "world 1: maxv" in lib.hl_last_error().decode()
    assert call(two, 2, verts=(0x10, 0)) == -1 and "world 1: verts" in lib.hl_last_error().decode()
    assert call(two, 2, nsteps=0) == 0  # nothing to do: no device work


@pytest.mark.parametrize("lattice,nz,rpitch,stride,force_stride,code,msg", [
    (27, 8, 80, 8, 0, 0, None),           # row-interleaved: accepted
    (27, 8, 8, 512, 0, 0, None),          # structure of arrays, explicit row pitch
    (27, 8, 80, 512, 0, -1, "stride == nx"),
    (27, 8, 24, 8, 0, -1, "rpitch"),
    (9, 1, 48, 8, 0, -2, "D3Q27"),
    (27, 8, 80, 8, 100, -1, "force stride"),
])
def test_moment_layouts_are_validated(lattice, nz, rpitch, stride, force_stride, code, msg):
    """rpitch selects structure of arrays (nx) or the row-interleaved layout
    (nfields * nx with stride = nx); anything else is rejected before any
    device work, as is a force stride shorter than the lattice."""
    lib = N.load()
    e = N.HLEnv()
    e.lattice, e.dtype, e.nx, e.ny, e.nz = lattice, 0, 8, 8, nz
    e.stride, e.rpitch, e.force_stride = stride, rpitch, force_stride
    e.moments[0] = 0x1000
    e.moments[1] = 0x2000
    e.nu = 0.1
    if force_stride:
        e.force = 0x4000
    # an accepted layout reaches the next check (cur), still before any device work
    rc = lib.hl_step(C.byref(e), 0 if code else 2, 1, 0, None, None)
    err = lib.hl_last_error().decode()
    if code:
        assert rc == code and msg in err, err
    else:
        assert rc == -1 and "cur must be" in err, err


def test_chain_layout_and_invalid_models_rejected_before_any_gpu_call():
    """hl_chain_create validates the model on the host before its device copy."""
    import numpy as np

    from openhomelbm_amd.chain import chain_from_mjcf

    lib = N.load()
    m = chain_from_mjcf(os.path.join(ROOT, "tests", "golden", "eel_2d.xml"))
    good = m.to_struct()
    assert C.sizeof(N.HLChainModel) == 3120  # 100 int32 + 340 double, no padding

    def rejected(edit, msg):
        s = m.to_struct()
        edit(s)
        h = C.c_void_p()
        assert lib.hl_chain_create(C.byref(s), C.byref(h)) == -1 and not h.value
        assert msg in lib.hl_last_error().decode()

    rejected(lambda s: setattr(s, "nbody", 0), "1..16 bodies")
    rejected(lambda s: setattr(s, "ndof", 17), "1..16 bodies")
    rejected(lambda s: setattr(s, "integrator", 5), "integrator")
    rejected(lambda s: setattr(s, "dt", 0.0), "dt must be positive")
    rejected(lambda s: s.parent.__setitem__(2, 3), "tree order")
    rejected(lambda s: s.dof_body.__setitem__(4, 0), "grouped by body")
    rejected(lambda s: s.dof_sign.__setitem__(4, 0.5), "hinge sign")
    rejected(lambda s: s.act_dof.__setitem__(0, 9), "actuator dof")
    assert good.nbody == 6 and good.ndof == 8 and np.isclose(good.gear[4], 0.2)
    with pytest.raises(NotImplementedError):
        import tempfile
        with tempfile.NamedTemporaryFile("w", suffix=".xml", delete=False) as f:
            f.write('<mujoco><worldbody><body><joint type="hinge" axis="1 0 0"/><geom size="0.1"/></body>'
                    '</worldbody></mujoco>')
        chain_from_mjcf(f.name)

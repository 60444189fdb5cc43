"""The C ABI library: it loads, exports every symbol include/avbd.h declares, and reports errors the reference's way
(codes + messages). No compute calls here: these run without a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "avbd.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(avbd_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_world_api():
    names = declared_symbols()
    for required in ["avbd_world_create", "avbd_batch_create", "avbd_world_step", "avbd_world_read_bodies",
                     "avbd_world_read_selected", "avbd_world_read_joints", "avbd_world_read_counters",
                     "avbd_world_read_pairs", "avbd_world_read_contacts", "avbd_world_write_state",
                     "avbd_world_adapt", "avbd_world_destroy", "avbd_last_error"]:
        assert required in names


def test_library_exports_every_declared_symbol(avbd):
    lib = avbd.load_library()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    # and the binding knows all of them
    from importlib import import_module
    native = import_module("babylon-avbd_amd._native")
    assert set(declared_symbols()) == set(native.SIGNATURES)


def test_errors_are_codes_with_messages(avbd):
    lib = avbd.load_library()
    native = __import__("importlib").import_module("babylon-avbd_amd._native")
    out = ctypes.c_void_p()
    rc = lib.avbd_world_create(None, None, ctypes.byref(out))
    assert rc == native.AVBD_ERR_INVALID
    assert b"scene" in lib.avbd_last_error()
    rc = lib.avbd_world_step(None, 1)
    assert rc == native.AVBD_ERR_DESTROYED
    assert b"destroyed" in lib.avbd_last_error()
    assert lib.avbd_version().startswith(b"avbd-mi355x")


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="a GPU is present")
def test_world_creation_without_a_gpu_fails_loudly(avbd):
    scene = avbd.configs.box_pile(counts=(2, 2, 2))
    with pytest.raises(avbd.AvbdError, match="no HIP device"):
        avbd.createSolverPipeline("hip").createSolver(scene)


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="a GPU is present")
def test_server_creation_without_a_gpu_fails_loudly(avbd):
    with pytest.raises(avbd.AvbdError, match="no HIP device"):
        avbd.serve.Server()


def test_scene_validation(avbd):
    s = avbd.AvbdScene3D()
    with pytest.raises(ValueError):
        s.addBox([1, -1, 1])
    with pytest.raises(ValueError):
        avbd.AvbdScene3D(up=(0, 2, 0))
    b = s.addBox([1, 1, 1], position=[0, 1, 0])
    with pytest.raises(ValueError):
        s.addJoint(b, b)
    with pytest.raises(ValueError):
        s.addSpring(None, b)
    with pytest.raises(ValueError):
        avbd.createSolverPipeline("webgpu")


def test_mass_properties_match_the_reference(avbd):
    m, moment, r = avbd.box_mass([1, 2, 3], 2)
    assert m == 12 and moment == [13, 10, 5]
    # capsule: cylinder + two hemispheres (capsuleShape.js:28-38) — closed forms
    import math
    m, (ir, ia, ir2), bound = avbd.capsule_mass(0.5, 2.0, 1.0)
    cyl = math.pi * 0.25 * 1.0
    caps = 4 / 3 * math.pi * 0.125
    assert abs(m - (cyl + caps)) < 1e-12 and ir == ir2 and bound == 1.0
    # a capsule of zero segment is a sphere: I = 2/5 m r²
    m, (ir, ia, _), _ = avbd.capsule_mass(0.5, 1.0, 1.0)
    assert abs(ia - 0.4 * m * 0.25) < 1e-12 and abs(ir - ia) < 1e-12


def test_solver_options_are_checked(avbd):
    """The solve kernel options the facade passes through the C ABI (solver, worldBlock, dataflowShare)."""
    opts = avbd.world._options
    assert opts({"solver": "worlds", "worldBlock": 512}).solver == 3
    assert opts({"solver": "worlds", "worldBlock": 512}).world_block == 512
    assert opts({"solver": "dataflow", "dataflowShare": 4}).dataflow_share == 4
    assert opts({}).solver == 0 and opts({}).world_block == 0 and opts({}).dataflow_share == 0
    with pytest.raises(ValueError, match="solver"):
        opts({"solver": "bogus"})
    with pytest.raises(ValueError, match="worldBlock"):
        opts({"worldBlock": 300})
    with pytest.raises(ValueError, match="dataflowShare"):
        opts({"dataflowShare": -1})

"""The JavaScript world API (babylon-avbd_amd/js: N-API addon + createSolverPipeline facade) against the Python host.

CPU: the addon loads under Node and exports one function per C entry point it binds; a scene built by the JS builder
packs to the same bytes as the same scene built by the Python builder (bodies, joints, joint info, ignored pairs,
filters); errors are thrown with the reference's messages, and world creation without a GPU fails loudly.
GPU: a world stepped through the JS facade ends bit-identical to the same world stepped through ctypes.
"""
import json
import math
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
CHECK = ROOT / "tests" / "js" / "facade_check.js"
NODE = shutil.which("node")
ADDON = ROOT / "babylon-avbd_amd" / "js" / "avbd_napi.node"

pytestmark = pytest.mark.skipif(NODE is None or not ADDON.exists(), reason="node or the built N-API addon is missing")


def run(mode, specs=None, tmp_path=None, *extra):
    args = [NODE, str(CHECK), mode]
    if specs is not None:
        f = tmp_path / "specs.json"
        f.write_text(json.dumps(specs))
        args.append(str(f))
    args += [str(x) for x in extra]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def build_py(avbd, spec):
    s = avbd.AvbdScene3D(**spec.get("params", {}))
    for kind, *a in spec["ops"]:
        if kind == "box":
            s.addBox(a[0], **(a[1] if len(a) > 1 else {}))
        elif kind == "sphere":
            s.addSphere(a[0], **(a[1] if len(a) > 1 else {}))
        elif kind == "capsule":
            s.addCapsule(a[0], a[1], **(a[2] if len(a) > 2 else {}))
        elif kind == "joint":
            s.addJoint(a[0], a[1], **(a[2] if len(a) > 2 else {}))
        elif kind == "spring":
            s.addSpring(a[0], a[1], **(a[2] if len(a) > 2 else {}))
        elif kind == "ignore":
            s.ignoreCollision(a[0], a[1])
        elif kind == "hull":
            s.addHull(a[0], **(a[1] if len(a) > 1 else {}))
        elif kind == "heightfield":
            s.addHeightfield(**a[0])
        elif kind == "staticMesh":
            s.addStaticMesh(**a[0])
    return s


def specs_mixed():
    rng = np.random.default_rng(11)
    ops = [["box", [20, 1, 20], {"density": 0, "friction": 0.5, "position": [0, -0.5, 0]}]]
    for k in range(12):
        q = rng.normal(size=4)
        ops.append(["box", list(rng.uniform(0.2, 1.5, 3)),
                    {"position": list(rng.uniform(-3, 3, 3)), "rotation": list(q), "density": float(rng.uniform(0.5, 3)),
                     "friction": float(rng.uniform(0.2, 0.9)), "velocity": list(rng.normal(size=3))}])
    ops.append(["sphere", 0.37, {"position": [1, 2, 3], "angularVelocity": [0.1, 0.2, 0.3]}])
    ops.append(["capsule", 0.2, 1.3, {"position": [0, 4, 0], "rotation": [0.1, 0.2, 0.3, 0.9], "density": 2.5}])
    ops.append(["box", [1, 2, 3], {"mass": 7.3, "group": 2, "collidesWith": 5}])
    ops.append(["joint", 1, 2, {"anchorA": [0.1, 0.2, 0.3], "anchorB": [-0.1, 0, 0.2], "angularStiffness": 50.0}])
    ops.append(["joint", None, 3, {"anchorA": [0, 5, 0], "breakForce": 40.0, "breakOnPull": True}])
    ops.append(["spring", 4, 5, {"anchorA": [0.2, 0, 0], "stiffness": 300.0}])
    ops.append(["spring", 5, 6, {"rest": 1.25}])
    ops.append(["ignore", 7, 8])
    cloud = [float(x) for x in rng.normal(size=60)]
    terrain = [["heightfield", {"heights": [0.1 * (k % 5) for k in range(20)], "columns": 5, "rows": 4,
                                "spacing": [1.5, 2.0], "origin": [-3, 0, -3], "friction": 0.8, "group": 3}],
               ["staticMesh", {"positions": [0, 0, 0, 4, 0, 0, 0, 1, 4, 4, 1, 4], "indices": [0, 2, 1, 1, 2, 3],
                               "thickness": 0.2}],
               ["hull", cloud, {"position": [1, 3, 0], "rotation": [0.2, 0.1, 0.3, 0.9], "density": 2.0}],
               ["hull", [0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], {"position": [-1, 2, 1], "mass": 3.0}]]
    return [{"params": {"gravity": -9.81, "iterations": 12, "timeStep": 1 / 120}, "ops": ops},
            {"ops": [["box", [1, 1, 1], {"position": [0, 0.5 * k, 0]}] for k in range(5)]},
            {"params": {"gravity": -9.81}, "ops": terrain}]


def test_addon_exports_the_bound_entry_points(tmp_path):
    out = run("exports")
    assert out["version"].startswith("avbd-mi355x")
    assert {"createWorld", "destroy", "step", "readBodies", "readSelected", "readJoints", "readCounters", "readPairs",
            "readContacts", "adapt", "writeBodies", "writeJoints", "stepTimed"} <= set(out["exports"])


def test_js_and_python_builders_pack_the_same_bytes(avbd, tmp_path):
    specs = specs_mixed()
    js = run("pack", specs, tmp_path)
    for spec, got in zip(specs, js):
        bodies, joints, info, ignore, filters = build_py(avbd, spec).packed()
        assert got["bodies"] == bodies.view(np.uint32).ravel().tolist()
        assert got["joints"] == joints.view(np.uint32).ravel().tolist()
        assert got["jointInfo"] == info.ravel().tolist()
        assert got["ignore"] == ignore.ravel().tolist()
        assert got["filters"] == (None if filters is None else filters.ravel().tolist())
        hulls = build_py(avbd, spec).hull_buffer()
        assert got["hulls"] == (hulls.ravel().tolist() if len(hulls) else None)


def test_errors_follow_the_reference(tmp_path):
    out = run("errors")
    assert out["backend"] == "Unsupported solver backend: webgpu"
    assert out["capsule"] == "Capsule height must be at least twice its positive radius"
    assert out["endpoints"] == "Constraint endpoints must be distinct bodies in this scene"
    if out["create"] != "has-gpu":
        assert "no HIP device" in out["create"]


@pytest.mark.gpu
def test_js_world_steps_like_the_python_world(avbd, tmp_path):
    rng = np.random.default_rng(5)
    ops = [["box", [30, 1, 30], {"density": 0, "friction": 0.5, "position": [0, -0.5, 0]}]]
    for k in range(60):
        ops.append(["box", list(rng.uniform(0.4, 1.2, 3)), {"position": [float(rng.uniform(-2, 2)), 1 + 1.3 * k / 4,
                                                                     float(rng.uniform(-2, 2))],
                                                        "rotation": list(rng.normal(size=4)), "friction": 0.5}])
    ops.append(["capsule", 0.25, 1.2, {"position": [0, 25, 0], "rotation": [0.3, 0.1, 0.2, 0.9]}])
    ops.append(["sphere", 0.4, {"position": [0.5, 27, 0.2]}])
    ops.append(["joint", 61, 62, {"anchorA": [0, 0.6, 0], "anchorB": [0, -0.4, 0]}])
    specs = [{"params": {"gravity": -9.81}, "ops": ops}]
    steps = 90
    js = run("step", specs, tmp_path, steps)[0]
    world = avbd.createSolverPipeline("hip").createSolver(build_py(avbd, specs[0]), dict(colorCap=32, colorRounds=16))
    world.step(steps)
    want = world.readBodies()
    assert js["bodies"] == want.view(np.uint32).ravel().tolist()
    assert js["pairs"] == world.readPairs().ravel().tolist()
    assert js["counters"]["pairs"] == world.readCounters()["pairs"]
    sel = world.readSelectedBodies([0, world.bodyCount - 1], posesOnly=True)
    assert js["poses"] == sel.view(np.uint32).ravel().tolist()
    assert js["contacts"] == world.readCounters()["contacts"]
    world.destroy()


@pytest.mark.gpu
def test_js_edits_land_like_the_python_edits(avbd, tmp_path):
    """The same edit calls through the JS facade and the Python facade, between steps: bit-identical bodies."""
    ops = [["box", [30, 1, 30], {"density": 0, "friction": 0.5, "position": [0, -0.5, 0]}]]
    for k in range(27):
        ops.append(["box", [1, 1, 1], {"position": [(k % 3) * 1.1 - 1.1, 0.5 + 1.05 * (k // 9), (k // 3 % 3) * 1.1 - 1.1]}])
    scene = {"params": {"gravity": -9.81}, "ops": ops}
    edits = [["setLinearVelocity", 3, [0, 4, 0]], ["applyImpulse", 5, [2, 0, 0], [0.2, 1.0, 0.1]],
             ["applyForce", 6, [0, 80, 0]], ["teleport", 7, [4, 3, 4], [0.1, 0.2, 0.3, 0.9]],
             ["editBodies", [{"body": 8, "position": [-4, 2, 0], "linearVelocity": [1, 0, 0], "torque": [0, 3, 0]},
                             {"body": 9, "group": 2, "collidesWith": 1}]],
             ["setPoses", [10, 11], [0, 6, 0, 1, 7, 1], [0, 0, 0, 1, 0.5, 0.5, 0.5, 0.5]],
             ["applyAngularImpulse", 12, [0, 0, 1]], ["setAngularVelocity", 13, [3, 0, 0]]]
    spec = {"scene": scene, "before": 20, "after": 30, "ops": edits}
    f = tmp_path / "spec.json"
    f.write_text(json.dumps(spec))
    r = subprocess.run([NODE, str(CHECK), "edits", str(f)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout)
    world = avbd.createSolverPipeline("hip").createSolver(build_py(avbd, scene), dict(colorCap=32, colorRounds=16))
    world.step(20)
    for method, *a in edits:
        if method == "setPoses":
            world.setPoses(np.array(a[0], np.uint32), np.array(a[1], np.float32), np.array(a[2], np.float32))
        else:
            getattr(world, method)(*a)
    world.step(30)
    assert js["bodies"] == world.readBodies().view(np.uint32).ravel().tolist()
    assert js["pairs"] == world.readPairs().ravel().tolist()
    world.destroy()


@pytest.mark.gpu
def test_reference_api_contracts_hold_on_the_hip_backend():
    """The reference's Babylon-free contracts (tests/backend-contract.js checkBackend/checkFloor,
    tests/packed-poses-contract.js), restated in tests/js/contract_check.js, pass through the N-API addon."""
    script = ROOT / "tests" / "js" / "contract_check.js"
    r = subprocess.run([NODE, str(script)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert out["backend"]["passed"] and out["floor"]["passed"] and out["restitution"]["reboundSpeed"] > 2
    assert out["sleeping"]["sleeping"] and out["sleeping"]["waking"]
    assert 0.35 < out["floor"]["floorHeight"] < 0.7
    assert all(out["packedPoses"].values())

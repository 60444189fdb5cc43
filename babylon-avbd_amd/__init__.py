"""MI355X-native AVBD 3D rigid-body step, behind the reference's world API.

    import importlib; avbd = importlib.import_module("babylon-avbd_amd")
    pipeline = avbd.createSolverPipeline(backend="hip")
    scene = avbd.AvbdScene3D(gravity=-9.81)
    ground = scene.addBox([100, 1, 100], density=0, position=[0, -0.5, 0])
    box = scene.addBox([1, 1, 1], position=[0, 3, 0])
    world = pipeline.createSolver(scene)
    world.step(); poses = world.readSelectedBodies([box], posesOnly=True)

The compute path is libavbd_hip.so (hand-written HIP kernels for gfx950 behind the C ABI of include/avbd.h); this
package is the host-side mirror of src/native/{pipeline,scenes,solver3D}.js.
"""
from ._native import AvbdError, load as load_library, LIB_PATH  # noqa: F401
from .scene import AvbdScene3D, box_mass, capsule_mass  # noqa: F401
from .world import HipBody, HipWorld, SolverPipeline, createSolverPipeline  # noqa: F401
from .hull import HullShape, MAX_HULL_VERTICES, convexHull, hullFromTriangles  # noqa: F401
from .terrain import cookHeightfield, cookStaticMesh, createTerrainBodies  # noqa: F401
from . import configs, serve, stats, tune  # noqa: F401

__all__ = ["AvbdScene3D", "HipBody", "HipWorld", "SolverPipeline", "createSolverPipeline", "AvbdError", "configs",
           "load_library", "HullShape", "MAX_HULL_VERTICES", "convexHull", "hullFromTriangles", "cookHeightfield",
           "cookStaticMesh", "createTerrainBodies"]

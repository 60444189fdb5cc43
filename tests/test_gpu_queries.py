"""Shape queries (src/gpu/shapeQueries.js) on the GPU against the oracle: every hit's normal, distance and body index
bit for bit, for rays and sphere casts into a pile of boxes, spheres and capsules (and one with convex hulls), with
group masks, membership and an ignore list; and the facade's result records (raycast, raycastAll, sphereCast) with the reference's validation."""
import numpy as np
import pytest

from oracle_lib import cast

pytestmark = pytest.mark.gpu


def pile(avbd, hulls=False):
    rng = np.random.default_rng(11)
    s = avbd.AvbdScene3D(gravity=-9.81)
    s.addBox([30, 1, 30], density=0, position=[0, -0.5, 0])
    shapes = [avbd.convexHull(rng.normal(size=(20, 3)) * rng.uniform(0.2, 0.5, 3)) for _ in range(3)] if hulls else []
    for k in range(300):
        pos = [float(rng.uniform(-6, 6)), float(0.6 + 0.3 * k / 10), float(rng.uniform(-6, 6))]
        if hulls and k % 4 == 3:
            s.addHullShape(shapes[k % 3], position=pos, rotation=list(rng.normal(size=4)))
        elif k % 3 == 0:
            s.addSphere(float(rng.uniform(0.25, 0.6)), position=pos)
        elif k % 3 == 1:
            s.addCapsule(float(rng.uniform(0.15, 0.35)), float(rng.uniform(0.8, 1.6)), position=pos,
                         rotation=list(rng.normal(size=4)))
        else:
            s.addBox(list(rng.uniform(0.3, 1.2, 3)), position=pos, rotation=list(rng.normal(size=4)))
    return s


def queries(n, radius, seed):
    rng = np.random.default_rng(seed)
    o = np.stack([rng.uniform(-9, 9, n), rng.uniform(0.2, 14, n), rng.uniform(-9, 9, n)], 1)
    target = np.stack([rng.uniform(-6, 6, n), rng.uniform(0, 4, n), rng.uniform(-6, 6, n)], 1)
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    md = np.where(rng.uniform(size=n) < 0.2, rng.uniform(1, 8, n), 1e3)
    q = np.concatenate([o, md[:, None], d, np.full((n, 1), radius)], 1).astype(np.float32)
    # the boundary normalises each direction in double precision (shapeQueries.js:159-163): the oracle gets the same
    d64 = q[:, 4:7].astype(np.float64)
    q[:, 4:7] = (d64 / np.sqrt(d64[:, 0] * d64[:, 0] + d64[:, 1] * d64[:, 1] + d64[:, 2] * d64[:, 2])[:, None])
    return q


@pytest.mark.parametrize("hulls", [False, True], ids=["boxes-spheres-capsules", "with-hulls"])
@pytest.mark.parametrize("radius", [0.0, 0.2])
def test_casts_match_the_oracle(avbd, radius, hulls):
    scene = pile(avbd, hulls)
    world = scene.createSolver(avbd.createSolverPipeline("hip"))
    world.step(60)
    n = world.bodyCount
    filters = np.full((n, 2), 0xFFFFFFFF, np.uint32)
    idx = np.arange(5, n, 7)
    filters[idx, 0], filters[idx, 1] = 2, 4
    world.setFilters(idx.tolist(), [2] * len(idx), [4] * len(idx))
    q = queries(2000, radius, 5 + int(radius * 10))
    # The oracle traces the directions the boundary traces: avbd_world_cast normalises each direction again (in
    # double precision), which moves about one direction in 2,000 by one ulp even though q is already normalised
    qo = q.copy()
    d64 = q[:, 4:7].astype(np.float64)
    qo[:, 4:7] = d64 / np.sqrt(d64[:, 0] * d64[:, 0] + d64[:, 1] * d64[:, 1] + d64[:, 2] * d64[:, 2])[:, None]
    bodies = world.readBodies()
    import ctypes
    from importlib import import_module
    native = import_module("babylon-avbd_amd._native")
    for opts in (dict(), dict(mask=1), dict(mask=0xFFFFFFFF, membership=4), dict(ignore=[1, 2, 3, 40])):
        want = cast(bodies, filters, qo, ignore=opts.get("ignore", ()), mask=opts.get("mask", 0xFFFFFFFF),
                    membership=opts.get("membership"), hulls=scene.hull_buffer())
        got = np.zeros_like(want)
        ign = np.ascontiguousarray(opts.get("ignore", []), np.uint32)
        o = native.QueryOptions(ign.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) if len(ign) else None, len(ign),
                                opts.get("mask", 0xFFFFFFFF), opts.get("membership") or 0,
                                int(opts.get("membership") is not None))
        native.check(world._lib.avbd_world_cast(world._h, native.ptr(q, ctypes.c_float), len(q), ctypes.byref(o),
                                                native.ptr(got, ctypes.c_float)))
        # every hit the same bits: body index, distance and normal
        gi, wi = got[:, 4].view(np.uint32), want[:, 4].view(np.uint32)
        assert np.array_equal(gi, wi), f"{opts}: hit bodies differ at {np.nonzero(gi != wi)[0][:5].tolist()}"
        hit = want[:, 3] >= 0
        assert np.array_equal(got[:, 3] >= 0, hit)
        assert np.allclose(got[hit, 3], want[hit, 3], rtol=1e-6, atol=0)
        assert np.abs(got[hit, 0:3] - want[hit, 0:3]).max(initial=0) <= 1e-6
        same = (got.view(np.uint32) == want.view(np.uint32)).all(axis=1)
        assert same.all(), f"{opts}: {int((~same).sum())} hits differ in bits, first at {np.flatnonzero(~same)[:5]}"
        assert (want[:, 3] >= 0).sum() > 500
    world.destroy()


def test_facade_records_and_validation(avbd):
    s = avbd.AvbdScene3D(gravity=0)
    s.addBox([2, 2, 2], position=[0, 0, 0])
    s.addSphere(0.5, position=[0, 0, 5])
    world = s.createSolver(avbd.createSolverPipeline("hip"))
    hit = world.raycast([0, 10, 0], [0, -2, 0])
    assert hit["index"] == 0 and hit["distance"] == 9 and hit["normal"] == [0, 1, 0]
    assert hit["center"] == [0, 1, 0] and hit["point"] == [0, 1, 0]
    cast_ = world.sphereCast([0, 0, 10], 0.5, [0, 0, -1])
    assert cast_["index"] == 1 and cast_["distance"] == 4 and cast_["point"] == [0, 0, 5.5]
    assert world.raycast([0, 10, 0], [0, -1, 0], {"maxDistance": 5}) is None
    assert world.raycast([0, 10, 0], [0, -1, 0], {"ignore": [0]}) is None
    assert [h and h["index"] for h in world.raycastAll([{"origin": [0, 0, 10], "direction": [0, 0, -1]},
                                                        {"origin": [5, 5, 5], "direction": [1, 0, 0]}])] == [1, None]
    for bad in ([[0, 0, 0], [0, 0, 0]], [[0, 0], [0, 1, 0]], [[0, float("nan"), 0], [0, 1, 0]]):
        with pytest.raises(ValueError):
            world.raycast(*bad)
    with pytest.raises(ValueError):
        world.raycast([0, 0, 0], [0, 1, 0], {"ignore": list(range(17))})
    with pytest.raises(ValueError):
        world.raycast([0, 0, 0], [0, 1, 0], {"includeTriggers": 1})
    world.destroy()

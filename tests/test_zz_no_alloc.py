"""avbd_world_cast_noalloc (the query entry point that allocates no device memory) against avbd_world_cast, through the
C ABI: two calls in a row on the same world, the second with fewer queries (the shared device array keeps the first
call's contents, which must not leak into the second), both bit-identical to the allocating entry point's hits; and
the size limit as an error that changes nothing."""
import ctypes
from importlib import import_module

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pile(avbd):
    rng = np.random.default_rng(3)
    s = avbd.AvbdScene3D(gravity=-9.81)
    s.addBox([30, 1, 30], density=0, position=[0, -0.5, 0])
    for k in range(200):
        pos = [float(rng.uniform(-5, 5)), float(0.6 + 0.25 * k / 10), float(rng.uniform(-5, 5))]
        if k % 3 == 0:
            s.addSphere(float(rng.uniform(0.25, 0.6)), position=pos)
        elif k % 3 == 1:
            s.addCapsule(float(rng.uniform(0.15, 0.35)), float(rng.uniform(0.8, 1.6)), position=pos,
                         rotation=list(rng.normal(size=4)))
        else:
            s.addBox(list(rng.uniform(0.3, 1.2, 3)), position=pos, rotation=list(rng.normal(size=4)))
    return s


def _queries(n, seed):
    """Raw query records (origin, maximum distance, direction, radius); directions left unnormalised on purpose: the
    entry points normalise them."""
    rng = np.random.default_rng(seed)
    o = np.stack([rng.uniform(-8, 8, n), rng.uniform(0.2, 12, n), rng.uniform(-8, 8, n)], 1)
    d = np.stack([rng.uniform(-5, 5, n), rng.uniform(0, 4, n), rng.uniform(-5, 5, n)], 1) - o
    d *= rng.uniform(0.5, 3.0, (n, 1))
    md = np.where(rng.uniform(size=n) < 0.2, rng.uniform(1, 8, n), 1e3)
    r = np.where(rng.uniform(size=n) < 0.5, 0.0, rng.uniform(0.05, 0.4, n))
    return np.ascontiguousarray(np.concatenate([o, md[:, None], d, r[:, None]], 1), np.float32)


def _call(lib, fn, h, q, opts):
    hits = np.full((len(q), 8), np.nan, np.float32)
    f32 = ctypes.POINTER(ctypes.c_float)
    rc = getattr(lib, fn)(h, q.ctypes.data_as(f32), len(q), ctypes.byref(opts) if opts is not None else None,
                          hits.ctypes.data_as(f32))
    return rc, hits


def test_noalloc_cast_matches_cast_twice(avbd):
    native = import_module("babylon-avbd_amd._native")
    lib = native.load()
    world = _pile(avbd).createSolver(avbd.createSolverPipeline("hip"))
    world.step(30)
    ign = np.array([1, 5, 9], np.uint32)
    masked = native.QueryOptions(ign.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(ign), 0xFFFFFFFF, 0, 0)
    for n, seed, opts in ((3000, 21, None), (700, 22, masked)):
        print(f"noalloc cast: {n} queries, seed {seed}", flush=True)
        q = _queries(n, seed)
        rc, want = _call(lib, "avbd_world_cast", world._h, q, opts)
        assert rc == 0, lib.avbd_last_error().decode()
        rc, got = _call(lib, "avbd_world_cast_noalloc", world._h, q, opts)
        assert rc == 0, lib.avbd_last_error().decode()
        assert (want[:, 3] >= 0).sum() > n // 4, "the queries hit something"
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), \
            f"{int((got.view(np.uint32) != want.view(np.uint32)).any(1).sum())} of {n} hits differ"
    world.destroy()


def test_noalloc_cast_rejects_too_many_queries(avbd):
    native = import_module("babylon-avbd_amd._native")
    lib = native.load()
    world = _pile(avbd).createSolver(avbd.createSolverPipeline("hip"))
    q = _queries(16385, 4)
    rc, hits = _call(lib, "avbd_world_cast_noalloc", world._h, q, None)
    assert rc == native.AVBD_ERR_INVALID
    assert "16384" in lib.avbd_last_error().decode()
    assert np.isnan(hits).all(), "a rejected call writes nothing"
    rc, hits = _call(lib, "avbd_world_cast_noalloc", world._h, q[:16384], None)
    assert rc == 0 and np.isfinite(hits[:, :4]).all()
    world.destroy()

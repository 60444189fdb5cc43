"""Seeded synthetic scenes of the benchmark workloads (boxes on a static ground, Y up).

Seeds are 0xA5BD0000 + workload number through numpy's PCG64. Choices the workload text leaves open are fixed here:
the ground box of the single large world is 240 x 1 x 240 m (its lattice spans 110 x 88 x 110 m), its lattice's lowest
layer sits 2 m above the ground, and its boxes have friction 0.5 like the 1,000-box pile.
"""
import math

import numpy as np

from .scene import AvbdScene3D

SEED = 0xA5BD0000


def _yaw_quaternions(rng, n):
    yaw = rng.uniform(0.0, 2.0 * math.pi, n)
    q = np.zeros((n, 4))
    q[:, 1] = np.sin(yaw / 2)
    q[:, 3] = np.cos(yaw / 2)
    return q


def _uniform_quaternions(rng, n):
    """Shoemake's uniform random unit quaternions."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    return np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2), b * np.sin(2 * np.pi * u3),
                     b * np.cos(2 * np.pi * u3)], axis=1)


def box_pile(n_side=10, spacing=1.1, centre=(0.0, 6.0, 0.0), seed=SEED + 0, friction=0.5, gravity=-9.81,
             iterations=10, dt=1 / 60, counts=None, ground=(100.0, 1.0, 100.0)):
    """Unit boxes on a jittered lattice above a static ground box (the 1,000-box pile: 10 x 10 x 10 at 1.1 m,
    centred 6 m up, jitter U(-0.05, 0.05) m, yaw U(0, 2π)). `counts` = (nx, ny, nz) overrides the cube."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = counts or (n_side, n_side, n_side)
    scene = AvbdScene3D(gravity=gravity, iterations=iterations, timeStep=dt)
    scene.addBox(list(ground), density=0, friction=friction, position=[0, -ground[1] / 2, 0])
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    grid = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    half = (np.array([nx, ny, nz], np.float64) - 1) / 2
    pos = (grid - half) * spacing + np.asarray(centre, np.float64)
    n = len(pos)
    pos += rng.uniform(-0.05, 0.05, (n, 3))
    rot = _yaw_quaternions(rng, n)
    scene.addBoxes(np.ones((n, 3)), pos, rot, density=1.0, friction=friction)
    return scene


def pile_1000(seed=SEED + 0):
    """The CPU-runnable workload: 1,000 unit boxes, 10 iterations, dt 1/60, friction 0.5, g = -9.81."""
    return box_pile(seed=seed)


def large_world(n=(50, 40, 50), spacing=2.2, seed=SEED + 1, iterations=10, friction=0.5):
    """The single large world: 100,000 boxes, half-extents U(0.25, 1.0) per axis, density 1, on a jittered
    50 x 40 x 50 lattice at 2.2 m above a static ground, uniform random orientations, v ~ N(0, 0.5²) per axis."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    scene = AvbdScene3D(gravity=-9.81, iterations=iterations, timeStep=1 / 60)
    extent = max(nx, nz) * spacing
    g = max(240.0, 2.2 * extent)
    scene.addBox([g, 1.0, g], density=0, friction=friction, position=[0, -0.5, 0])
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    grid = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    count = len(grid)
    pos = np.empty((count, 3))
    pos[:, 0] = (grid[:, 0] - (nx - 1) / 2) * spacing
    pos[:, 1] = 2.0 + grid[:, 1] * spacing
    pos[:, 2] = (grid[:, 2] - (nz - 1) / 2) * spacing
    pos += rng.uniform(-0.05, 0.05, (count, 3))
    half = rng.uniform(0.25, 1.0, (count, 3))
    rot = _uniform_quaternions(rng, count)
    vel = rng.normal(0.0, 0.5, (count, 3))
    scene.addBoxes(2 * half, pos, rot, density=1.0, friction=friction, velocities=vel)
    return scene


def pile_of(n, seed, friction=0.5, iterations=10):
    """A pile of n unit boxes like the 1,000-box pile: the smallest near-cube lattice holding n, lowest layer 1.05 m
    above the ground."""
    side = max(1, int(math.ceil(n ** (1 / 3))))
    nx = side
    nz = side
    ny = int(math.ceil(n / (nx * nz)))
    rng = np.random.default_rng(seed)
    scene = AvbdScene3D(gravity=-9.81, iterations=iterations, timeStep=1 / 60)
    scene.addBox([100, 1, 100], density=0, friction=friction, position=[0, -0.5, 0])
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    grid = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)[:n]
    pos = np.empty((n, 3))
    pos[:, 0] = (grid[:, 0] - (nx - 1) / 2) * 1.1
    pos[:, 1] = 1.05 + grid[:, 1] * 1.1
    pos[:, 2] = (grid[:, 2] - (nz - 1) / 2) * 1.1
    pos += rng.uniform(-0.05, 0.05, (n, 3))
    scene.addBoxes(np.ones((n, 3)), pos, _yaw_quaternions(rng, n), density=1.0, friction=friction)
    return scene


def batch_sizes(count=4096, lo=8, hi=2048, seed=SEED + 3):
    """Body counts log-uniform in [lo, hi] and frictions U(0.2, 0.9) of the batched-worlds workload."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), count))).astype(int)
    frictions = rng.uniform(0.2, 0.9, count)
    return sizes, frictions


def batched_piles(count=4096, lo=8, hi=2048, seed=SEED + 3):
    sizes, frictions = batch_sizes(count, lo, hi, seed)
    return [pile_of(int(n), seed + 1000 + k, friction=float(f)) for k, (n, f) in enumerate(zip(sizes, frictions))]

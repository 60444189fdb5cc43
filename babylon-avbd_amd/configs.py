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


def jointed_mix(n=64, spheres=2000, capsules=20000, area=40.0, seed=SEED + 2, iterations=20, capsule_grid=None):
    """The jointed + contact mix: an n x n net of 0.2 m boxes joined to their 4-neighbours by hard ball joints at the
    shared face centres (anchors ±0.1 m; the penalty ramp stiffens them), dropped onto `spheres` static spheres
    (radius U(0.3, 1.0), centres uniform over an area x area square on the ground plane) and `capsules` free capsules
    (radius 0.2, segment half-length 0.4, so 1.2 m tip to tip). dt 1/120, 20 iterations, g = -9.81.

    Choices the workload text leaves open: a static ground box (area + 20 m square, 1 m thick, top at y = 0); the
    capsules start on a jittered lattice (25 x 32 x 25 for 20,000: 1.6 m apart across the area, 1.3 m apart upwards
    from y = 2.5 m) with uniform random orientations; the net starts flat 3 m above the highest capsule."""
    rng = np.random.default_rng(seed)
    scene = AvbdScene3D(gravity=-9.81, iterations=iterations, timeStep=1 / 120)
    g = area + 20.0
    scene.addBox([g, 1.0, g], density=0, friction=0.5, position=[0, -0.5, 0])
    radii = rng.uniform(0.3, 1.0, spheres)
    xz = rng.uniform(-area / 2, area / 2, (spheres, 2))
    for r, (x, z) in zip(radii, xz):
        scene.addSphere(float(r), density=0, friction=0.5, position=[float(x), 0.0, float(z)])
    if capsule_grid is None:
        side = max(1, int(round(math.sqrt(capsules / 32)))) if capsules else 1
        capsule_grid = (side, int(math.ceil(capsules / (side * side))) if capsules else 0, side)
    nx, ny, nz = capsule_grid
    pitch = area / max(nx, 1)
    placed = 0
    top = 0.0
    rot = _uniform_quaternions(rng, capsules)
    jitter = rng.uniform(-0.05, 0.05, (capsules, 3))
    for iy in range(ny):
        for ix in range(nx):
            for iz in range(nz):
                if placed == capsules:
                    break
                p = [(ix - (nx - 1) / 2) * pitch, 2.5 + iy * 1.3, (iz - (nz - 1) / 2) * pitch]
                p = [p[k] + float(jitter[placed, k]) for k in range(3)]
                scene.addCapsule(0.2, 1.2, friction=0.5, position=p, rotation=list(rot[placed]))
                top = max(top, p[1])
                placed += 1
    first = scene.body_count
    y = top + 3.0
    half = (n - 1) / 2
    for i in range(n):
        for j in range(n):
            scene.addBox([0.2, 0.2, 0.2], friction=0.5, position=[(i - half) * 0.2, y, (j - half) * 0.2])
    for i in range(n):
        for j in range(n):
            k = first + i * n + j
            if i + 1 < n:
                scene.addJoint(k, k + n, anchorA=[0.1, 0, 0], anchorB=[-0.1, 0, 0])
            if j + 1 < n:
                scene.addJoint(k, k + 1, anchorA=[0, 0, 0.1], anchorB=[0, 0, -0.1])
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


def boxes_world(n, seed, iterations=10, friction=0.5, spacing=2.2):
    """n boxes with the single large world's distributions (half-extents U(0.25, 1.0), uniform random orientations,
    v ~ N(0, 0.5²) per axis) on a jittered near-cubic lattice at 2.2 m, lowest layer 2 m above a static ground."""
    rng = np.random.default_rng(seed)
    side = max(1, int(math.ceil(n ** (1 / 3))))
    nx = nz = side
    ny = int(math.ceil(n / (nx * nz)))
    scene = AvbdScene3D(gravity=-9.81, iterations=iterations, timeStep=1 / 60)
    g = max(40.0, 2.2 * side * spacing)
    scene.addBox([g, 1.0, g], density=0, friction=friction, position=[0, -0.5, 0])
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    grid = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)[:n]
    pos = np.empty((n, 3))
    pos[:, 0] = (grid[:, 0] - (nx - 1) / 2) * spacing
    pos[:, 1] = 2.0 + grid[:, 1] * spacing
    pos[:, 2] = (grid[:, 2] - (nz - 1) / 2) * spacing
    pos += rng.uniform(-0.05, 0.05, (n, 3))
    half = rng.uniform(0.25, 1.0, (n, 3))
    scene.addBoxes(2 * half, pos, _uniform_quaternions(rng, n), density=1.0, friction=friction,
                   velocities=rng.normal(0.0, 0.5, (n, 3)))
    return scene


def serving_pool(count=64, lo=16, hi=50000, seed=SEED + 4):
    """World sizes of the serving workload: log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), count))).astype(int)


def serving_requests(count, rate=2000.0, worlds=64, max_steps=120, seed=SEED + 5):
    """Poisson arrivals (exponential gaps at `rate` per second) of rollout requests: (arrival s, world, steps U{1..120})."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    return t, rng.integers(0, worlds, count), rng.integers(1, max_steps + 1, count)

"""Pins of the oracle's analytic capsule contacts (oracle/collide.hpp capsule section; reference
src/gpu/capsuleCollision.wgsl:1-84 and capsuleShape.js:4-41) by checks that do not reuse its formulas (CPU only):

* mass and inertia of a capsule against numerical integration of the solid;
* capsule/sphere, capsule/capsule and capsule/box: the contact depth against an independent distance (a bounded
  minimisation over the segment parameters and the box's local coordinates), both contact points on their surfaces,
  the normal along the closest-point direction, and the same contacts (flipped) with the arguments swapped;
* a segment inside the box: moving the capsule out along the contact normal by the reported depth separates it;
* a capsule lying on a box face: two end-cap points (features 4<<24 | 0 and | 1) at equal depth;
* a dropped capsule comes to lie on the ground on two end-cap contacts with its axis at r minus the collision margin.
"""
import ctypes
import math

import numpy as np
import pytest
from scipy.optimize import minimize

from oracle_lib import OracleWorld, lib
from test_oracle_gpusem import quat_to_mat, random_quat

D = ctypes.POINTER(ctypes.c_double)
U = ctypes.POINTER(ctypes.c_uint32)
F = ctypes.POINTER(ctypes.c_float)
CAPSULE = 4 << 24


def record(avbd, kind, dims, position, rotation):
    s = avbd.AvbdScene3D()
    if kind == "capsule":
        s.addCapsule(*dims)
    elif kind == "sphere":
        s.addSphere(dims[0])
    else:
        s.addBox(list(dims))
    rec = s.packed()[0][-1].copy()
    rec[0:3] = position
    rec[4:8] = rotation
    return rec


def pair_contacts(recA, recB, precision=64):
    a = np.ascontiguousarray(recA, np.float32)
    b = np.ascontiguousarray(recB, np.float32)
    feats = np.zeros(8, np.uint32)
    pts = np.zeros((8, 6))
    normal = np.zeros(3)
    n = lib().gs_probe_pair_contacts(precision, a.ctypes.data_as(F), b.ctypes.data_as(F), 1, feats.ctypes.data_as(U),
                                     pts.ctypes.data_as(D), normal.ctypes.data_as(D))
    return feats[:n], pts[:n], normal


def segment(rec):
    """End points and radius of a capsule record (float64)."""
    r = float(rec[16]) / 2
    half = max(0.0, float(rec[17]) / 2 - r)
    axis = quat_to_mat(rec[4:8].astype(np.float64)) @ np.array([0.0, half, 0.0])
    c = rec[0:3].astype(np.float64)
    return c - axis, c + axis, r


def seg_point_distance(a, b, p):
    res = minimize(lambda t: np.sum((a + t[0] * (b - a) - p) ** 2), [0.5], bounds=[(0, 1)], method="L-BFGS-B",
                   options=dict(ftol=1e-15, gtol=1e-12))
    return math.sqrt(max(res.fun, 0.0))


def seg_seg_distance(a, b, c, d):
    best = math.inf
    for start in ([0.5, 0.5], [0.0, 0.0], [1.0, 1.0], [0.0, 1.0], [1.0, 0.0]):
        res = minimize(lambda x: np.sum((a + x[0] * (b - a) - c - x[1] * (d - c)) ** 2), start,
                       bounds=[(0, 1), (0, 1)], method="L-BFGS-B", options=dict(ftol=1e-15, gtol=1e-12))
        best = min(best, res.fun)
    return math.sqrt(max(best, 0.0))


def seg_box_distance(a, b, centre, q, h):
    """Distance from segment ab to the solid box (centre, rotation q, half extents h): minimise over (t, box point)."""
    R = quat_to_mat(q)
    la, lb = R.T @ (a - centre), R.T @ (b - centre)
    best = math.inf
    for t0 in (0.0, 0.25, 0.5, 0.75, 1.0):
        p0 = np.clip(la + t0 * (lb - la), -h, h)
        res = minimize(lambda x: np.sum((la + x[0] * (lb - la) - x[1:]) ** 2), [t0, *p0],
                       bounds=[(0, 1)] + [(-hk, hk) for hk in h], method="L-BFGS-B",
                       options=dict(ftol=1e-16, gtol=1e-13))
        best = min(best, res.fun)
    return math.sqrt(max(best, 0.0))


def box_surface_distance(p, centre, q, h):
    """Signed distance from p to the box surface (negative inside)."""
    x = quat_to_mat(q).T @ (p - centre)
    outside = np.maximum(np.abs(x) - h, 0.0)
    if (outside > 0).any():
        return float(np.linalg.norm(outside))
    return -float(np.min(h - np.abs(x)))


def test_capsule_mass_properties_match_integration(avbd):
    r, height, density = 0.3, 1.4, 2.0
    s = avbd.AvbdScene3D()
    s.addCapsule(r, height, density=density)
    rec = s.packed()[0][-1]
    m, moment = float(rec[19]), rec[20:23].astype(np.float64)
    half = height / 2 - r
    n = 160
    xs = (np.arange(n) + 0.5) / n * 2 * r - r
    ys = (np.arange(n) + 0.5) / n * height - height / 2
    X, Y, Z = np.meshgrid(xs, ys, xs, indexing="ij")
    dy = np.maximum(np.abs(Y) - half, 0.0)
    inside = X * X + dy * dy + Z * Z <= r * r
    cell = (2 * r / n) * (height / n) * (2 * r / n) * density
    mass = inside.sum() * cell
    iyy = ((X * X + Z * Z) * inside).sum() * cell
    ixx = ((Y * Y + Z * Z) * inside).sum() * cell
    assert m == pytest.approx(mass, rel=3e-3)
    assert moment[1] == pytest.approx(iyy, rel=5e-3)
    assert moment[0] == pytest.approx(ixx, rel=5e-3) and moment[2] == pytest.approx(ixx, rel=5e-3)
    assert float(rec[23]) == pytest.approx(height / 2)
    assert rec[39] == -1.0


def _check_swap(recA, recB):
    fa, pa, na = pair_contacts(recA, recB)
    fb, pb, nb = pair_contacts(recB, recA)
    assert len(fa) == len(fb)
    assert np.array_equal(np.sort(fa), np.sort(fb))
    np.testing.assert_allclose(na, -nb, atol=1e-12)
    np.testing.assert_allclose(np.sort(pa[:, :3], axis=0), np.sort(pb[:, 3:], axis=0), atol=1e-12)
    return fa, pa, na


@pytest.mark.parametrize("seed", range(12))
def test_capsule_sphere_depth_and_points(avbd, seed):
    rng = np.random.default_rng(seed)
    cap = record(avbd, "capsule", (rng.uniform(0.1, 0.5), rng.uniform(1.2, 3.0)), rng.uniform(-1, 1, 3),
                 random_quat(rng))
    a, b, ra = segment(cap)
    rs = rng.uniform(0.1, 0.8)
    target = a + rng.uniform(-0.2, 1.2) * (b - a)
    direction = rng.normal(size=3)
    direction /= np.linalg.norm(direction)
    centre = target + direction * (ra + rs) * rng.uniform(0.5, 1.2)
    sph = record(avbd, "sphere", (rs,), centre, (0, 0, 0, 1))
    centre = sph[0:3].astype(np.float64)
    rs = float(sph[16]) / 2
    dist = seg_point_distance(a, b, centre)
    feats, pts, normal = _check_swap(cap, sph)
    if dist > ra + rs + 1e-9:
        assert len(feats) == 0
        return
    assert list(feats) == [CAPSULE]
    xA, xB = pts[0, :3], pts[0, 3:]
    assert seg_point_distance(a, b, xA) == pytest.approx(ra, abs=1e-7)
    assert np.linalg.norm(xB - centre) == pytest.approx(rs, abs=1e-9)
    depth = float(np.dot(xA - xB, -normal))  # normal points from B (sphere) towards A (capsule)
    assert depth == pytest.approx(ra + rs - dist, abs=1e-7)


@pytest.mark.parametrize("seed", range(16))
def test_capsule_capsule_depth_and_points(avbd, seed):
    rng = np.random.default_rng(100 + seed)
    A = record(avbd, "capsule", (rng.uniform(0.1, 0.4), rng.uniform(1.0, 3.0)), rng.uniform(-1, 1, 3),
               random_quat(rng))
    a0, a1, ra = segment(A)
    rb = rng.uniform(0.1, 0.4)
    qb = random_quat(rng)
    direction = rng.normal(size=3)
    direction /= np.linalg.norm(direction)
    anchor = a0 + rng.uniform(0, 1) * (a1 - a0)
    B = record(avbd, "capsule", (rb, rng.uniform(1.0, 3.0)), anchor + direction * rng.uniform(0.1, 1.5), qb)
    b0, b1, rb = segment(B)
    dist = seg_seg_distance(a0, a1, b0, b1)
    feats, pts, normal = _check_swap(A, B)
    if dist > ra + rb + 1e-9:
        assert len(feats) == 0
        return
    assert len(feats) >= 1 and all((f >> 24) == 4 for f in feats)
    for xA, xB in zip(pts[:, :3], pts[:, 3:]):
        assert seg_point_distance(a0, a1, xA) == pytest.approx(ra, abs=1e-7)
        assert seg_point_distance(b0, b1, xB) == pytest.approx(rb, abs=1e-7)
    if len(feats) == 1 and dist > 1e-6:
        depth = float(np.dot(pts[0, :3] - pts[0, 3:], -normal))
        assert depth == pytest.approx(ra + rb - dist, abs=1e-7)


def test_parallel_capsules_give_two_points(avbd):
    A = record(avbd, "capsule", (0.2, 2.0), (0, 0, 0), (0, 0, 0, 1))
    B = record(avbd, "capsule", (0.2, 2.0), (0.35, 0.5, 0), (0, 0, 0, 1))
    feats, pts, normal = _check_swap(A, B)
    assert sorted(feats) == [CAPSULE, CAPSULE | 1]
    np.testing.assert_allclose(normal, [-1, 0, 0], atol=1e-12)
    depths = [float(np.dot(p[:3] - p[3:], -normal)) for p in pts]
    np.testing.assert_allclose(depths, [0.05, 0.05], atol=1e-7)


@pytest.mark.parametrize("seed", range(16))
def test_capsule_box_outside_depth_and_points(avbd, seed):
    rng = np.random.default_rng(200 + seed)
    h = rng.uniform(0.2, 1.0, 3)
    qb = random_quat(rng)
    box = record(avbd, "box", tuple(2 * h), (0, 0, 0), qb)
    h = box[16:19].astype(np.float64) / 2
    qb = box[4:8].astype(np.float64)
    for attempt in range(50):  # a capsule whose segment stays outside the box (the inside case is tested below)
        r = rng.uniform(0.1, 0.4)
        direction = rng.normal(size=3)
        direction /= np.linalg.norm(direction)
        centre = direction * (np.linalg.norm(h) + r * rng.uniform(-0.5, 1.5))
        cap = record(avbd, "capsule", (r, rng.uniform(1.0, 2.5)), centre, random_quat(rng))
        a, b, r = segment(cap)
        dist = seg_box_distance(a, b, np.zeros(3), qb, h)
        if dist > 1e-4:
            break
    feats, pts, normal = _check_swap(cap, box)
    if dist > r + 1e-9:
        assert len(feats) == 0
        return
    assert len(feats) >= 1 and all((f >> 24) == 4 for f in feats)
    for xA, xB in zip(pts[:, :3], pts[:, 3:]):
        assert seg_point_distance(a, b, xA) == pytest.approx(r, abs=2e-6)
        assert box_surface_distance(xB, np.zeros(3), qb, h) == pytest.approx(0.0, abs=5e-7)
    if len(feats) == 1:
        depth = float(np.dot(pts[0, :3] - pts[0, 3:], -normal))
        assert depth == pytest.approx(r - dist, abs=2e-6)


@pytest.mark.parametrize("seed", range(10))
def test_capsule_box_inside_separates_along_the_normal(avbd, seed):
    rng = np.random.default_rng(300 + seed)
    h = rng.uniform(0.4, 1.0, 3)
    box = record(avbd, "box", tuple(2 * h), (0, 0, 0), random_quat(rng))
    h = box[16:19].astype(np.float64) / 2
    qb = box[4:8].astype(np.float64)
    cap = record(avbd, "capsule", (0.15, rng.uniform(0.6, 1.2)), rng.uniform(-0.2, 0.2, 3), random_quat(rng))
    a, b, r = segment(cap)
    assert seg_box_distance(a, b, np.zeros(3), qb, h) < 1e-9
    feats, pts, normal = _check_swap(cap, box)
    assert len(feats) >= 1
    depths = [float(np.dot(p[:3] - p[3:], -normal)) for p in pts]
    assert max(depths) > r
    # the normal points from the box towards the capsule: moving the capsule out along it by the deepest point's
    # depth leaves the segment at least r from the box (SAT axis, or both end caps past the face plane)
    shift = normal * (max(depths) + 1e-6)
    assert seg_box_distance(a + shift, b + shift, np.zeros(3), qb, h) >= r - 1e-5


def test_capsule_lying_on_a_box_face_gives_two_end_points(avbd):
    box = record(avbd, "box", (4, 1, 4), (0, -0.5, 0), (0, 0, 0, 1))
    s = math.sin(math.pi / 4)
    cap = record(avbd, "capsule", (0.2, 1.2), (0.1, 0.19, -0.2), (0, 0, s, s))  # axis along -x, 1 cm deep
    feats, pts, normal = _check_swap(cap, box)
    assert sorted(feats) == [CAPSULE, CAPSULE | 1]
    np.testing.assert_allclose(normal, [0, 1, 0], atol=1e-7)
    depths = [float(np.dot(p[:3] - p[3:], -normal)) for p in pts]
    np.testing.assert_allclose(depths, [0.01, 0.01], atol=1e-6)


def test_dropped_capsule_rests_lying_at_radius_minus_margin(avbd):
    s = avbd.AvbdScene3D(gravity=-9.81)
    s.addBox([20, 1, 20], density=0, friction=0.5, position=[0, -0.5, 0])
    q = np.array([0.2, 0.1, 0.5, 0.8])
    s.addCapsule(0.25, 1.5, friction=0.5, position=[0, 2.0, 0], rotation=list(q / np.linalg.norm(q)))
    w = OracleWorld([s], precision=32)
    w.step(600)
    b = w.bodies()[1]
    axis = quat_to_mat(b[4:8]) @ np.array([0.0, 1.0, 0.0])
    assert abs(axis[1]) < 0.02, "the capsule lies down"
    assert w.counts()["contacts"] == 2
    assert b[1] == pytest.approx(0.25 - 0.01, abs=1e-3)

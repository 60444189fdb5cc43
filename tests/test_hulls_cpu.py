"""Convex hulls and terrain on the CPU: the cooking in the library (avbd_hull_from_points / avbd_hull_from_triangles,
no device touched), the scene and terrain builders, and the oracle's hull contacts.

Cooking is checked against the reference's own tests (reference/three-avbd/tests/hull.test.ts:15-150: a box's faces,
edges, volume, moments and radius; planar faces with every vertex behind them and Euler's V - E + F = 2 on a random
cloud; the principal frame mapping back onto the input; no hull for flat or tiny sets; a cut cube's volume and seven
faces; sphere points simplified to 32 vertices; a cube from triangles; an open surface) and against what the
mathematics fixes: tetrahedron volumes and centroids, box inertia, every input point inside its hull.

The oracle's hull contacts (oracle/hull.hpp, a sequential restatement of wgsl-hull.ts and capsulePoly) are pinned by
independent computations: the best separating axis equals a brute-force search over every face normal and every edge
pair's cross product with support-function separations; separated pairs make no contact; contact points lie on both
surfaces; a hull cooked from a box's corners collides like the box/box SAT (which the golden trajectories pin); a
sphere's and a capsule's contacts are at the distances a constrained minimisation (scipy) finds.
"""
import importlib
import math

import numpy as np
import pytest
from scipy.optimize import minimize

from oracle_lib import OracleWorld, probe_hull_pair

avbd = importlib.import_module("babylon-avbd_amd")


def box_points(w, h, d):
    return [c for x in (-w / 2, w / 2) for y in (-h / 2, h / 2) for z in (-d / 2, d / 2) for c in (x, y, z)]


def lcg(seed):
    s = seed

    def rnd():
        nonlocal s
        s = (s * 16807) % 2147483647
        return s / 2147483647
    return rnd


def quat_rotate(q, v):
    x, y, z, w = q
    u = np.array([x, y, z])
    t = 2 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


# --- cooking: the reference's known answers ---------------------------------------------------------------------------

def test_box_faces_edges_volume_moments():
    h = avbd.convexHull(box_points(2, 1, 0.5) + [0, 0, 0, 0.3, 0.1, 0.1])
    assert len(h.faces) == 6 and all(len(f["verts"]) == 4 for f in h.faces)
    assert len(h.edges) == 12 and h.vertices.shape == (8, 3)
    assert abs(h.volume - 1) < 1e-12
    m = sorted(h.moments)
    assert abs(m[0] - 1.25 / 12) < 1e-9 and abs(m[1] - 4.25 / 12) < 1e-9 and abs(m[2] - 5 / 12) < 1e-9
    assert abs(h.radius - math.hypot(1, 0.5, 0.25)) < 1e-6


def test_random_cloud_planar_faces_and_euler():
    rnd = lcg(7)
    pts = []
    for _ in range(200):
        pts += [(rnd() * 2 - 1) * 1.5 + 3, (rnd() * 2 - 1) * 0.7 - 2, rnd() * 2 - 1]
    h = avbd.convexHull(pts)
    for f in h.faces:
        n = np.array(f["normal"])
        assert abs(np.linalg.norm(n) - 1) < 1e-5
        assert np.all(h.vertices @ n <= f["d"] + 1e-5)
    assert len(h.vertices) - len(h.edges) + len(h.faces) == 2
    for a, b, f, g in h.edges:
        assert a in h.faces[f]["verts"] and b in h.faces[f]["verts"]
        assert a in h.faces[g]["verts"] and b in h.faces[g]["verts"]


def test_principal_frame_maps_back():
    angle = math.pi / 6
    c, s = math.cos(angle), math.sin(angle)
    src = np.array(box_points(3, 1, 0.4)).reshape(-1, 3)
    turned = np.stack([c * src[:, 0] - s * src[:, 1] + 1, s * src[:, 0] + c * src[:, 1] - 2, src[:, 2] + 0.5], 1)
    h = avbd.convexHull(turned)
    assert np.allclose(h.center, [1, -2, 0.5], atol=1e-6)
    assert np.allclose(sorted(h.size), [0.4, 1, 3], atol=1e-5)
    for v in h.vertices.astype(np.float64):
        p = quat_rotate(h.rotation, v) + h.center
        assert np.min(np.linalg.norm(turned - p, axis=1)) < 1e-5


def test_flat_and_tiny_sets_have_no_hull():
    assert avbd.convexHull([0, 0, 0, 1, 0, 0, 0, 1, 0]) is None
    assert avbd.convexHull([0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0, 0.5, 0.5, 0]) is None


def test_cut_cube_closed_with_seven_faces():
    rnd = lcg(11)
    pts = []
    for k in range(400):
        face = k % 4
        u, v = rnd(), rnd()
        p = [0, u, v] if face == 0 else [u, 0, v] if face == 1 else [u, v, 0] if face == 2 else [u, v, 1.6 - u - v]
        if (0 <= p[2] <= 1) if face == 3 else (sum(p) <= 1.6):
            pts += p
    for c in ([0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0.6, 0], [0.6, 1, 0], [1, 0, 0.6], [0, 1, 0.6],
              [0.6, 0, 1], [0, 0.6, 1]):
        pts += c
    h = avbd.convexHull(pts)
    assert len(h.vertices) - len(h.edges) + len(h.faces) == 2
    assert len(h.faces) == 7
    assert abs(h.volume - (1 - (1.4 ** 3 - 3 * 0.4 ** 3) / 6)) < 1e-4


def test_sphere_points_simplified_and_triangles():
    rnd = lcg(3)
    last = 0
    for n in (100, 400):
        pts = []
        for _ in range(n):
            u = rnd() * 2 - 1
            a = rnd() * 2 * math.pi
            r = math.sqrt(1 - u * u)
            pts += [r * math.cos(a), r * math.sin(a), u]
        h = avbd.convexHull(pts)
        assert len(h.vertices) == min(n, avbd.MAX_HULL_VERTICES)
        assert len(h.faces) <= 255 and len(h.edges) <= 255
        assert len(h.vertices) - len(h.edges) + len(h.faces) == 2
        assert max(last * 0.95, 3) < h.volume < 4 / 3 * math.pi
        last = h.volume
    cube = avbd.hullFromTriangles(
        [-1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, -1, 1, 1, -1, 1, 1, 1, 1, -1, 1, 1],
        [0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4, 2, 3, 7, 2, 7, 6, 1, 2, 6, 1, 6, 5, 3, 0, 4, 3, 4, 7])
    assert len(cube.faces) == 6 and abs(cube.volume - 8) < 1e-12
    assert avbd.hullFromTriangles([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], [0, 2, 1, 0, 1, 3, 0, 3, 2]) is None


# --- cooking: closed forms ----------------------------------------------------------------------------------------------

@pytest.mark.parametrize("seed", range(5))
def test_tetrahedron_volume_centroid_and_points_inside(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(4, 3)) * [1, 2, 0.5] + rng.normal(size=3) * 3
    h = avbd.convexHull(p)
    assert len(h.faces) == 4 and len(h.edges) == 6
    vol = abs(np.linalg.det(p[1:] - p[0])) / 6
    assert abs(h.volume - vol) < 1e-9 * max(1, vol)
    assert np.allclose(h.center, p.mean(0), atol=1e-9)
    # A cloud's hull holds every input point, and its vertices are input points
    cloud = rng.normal(size=(60, 3)) * [1, 0.5, 2]
    g = avbd.convexHull(cloud)
    for f in g.faces:
        n = np.array(f["normal"])
        local = np.array([quat_rotate([-x for x in g.rotation[:3]] + [g.rotation[3]], q - np.array(g.center))
                          for q in cloud])
        assert np.all(local @ n <= f["d"] + 1e-6)
    world = np.array([quat_rotate(g.rotation, v) + g.center for v in g.vertices.astype(np.float64)])
    assert all(np.min(np.linalg.norm(cloud - v, axis=1)) < 1e-5 for v in world)


def test_box_inertia_about_turned_axes():
    h = avbd.convexHull(box_points(1, 2, 3))
    assert abs(h.volume - 6) < 1e-12
    # principal moments per unit density V (b² + c²)/12
    assert np.allclose(sorted(h.moments), sorted([6 * (4 + 9) / 12, 6 * (1 + 9) / 12, 6 * (1 + 4) / 12]), rtol=1e-12)


# --- scene and terrain builders -----------------------------------------------------------------------------------------

def test_add_hull_records_and_shared_buffer():
    s = avbd.AvbdScene3D()
    shape = avbd.convexHull(box_points(1, 1, 1))
    a = s.addHullShape(shape, density=2.0, position=[0, 1, 0])
    b = s.addHullShape(shape, density=0, position=[3, 1, 0])
    c = s.addHull([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], friction=0.3)
    ra, rb, rc = s.record(a), s.record(b), s.record(c)
    assert ra[39] == 3.0 and rb[39] == 3.0 and rc[39] == 3.0 + shape.words
    assert abs(ra[19] - 2.0) < 1e-6 and rb[19] == 0 and np.allclose(ra[20:23], np.array(shape.moments) * 2)
    assert np.allclose(ra[16:19], shape.size) and abs(ra[23] - shape.radius) < 1e-6
    buf = s.hull_buffer()
    assert len(buf) == shape.words + avbd.convexHull([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1]).words
    assert buf[shape.words, 0] == shape.words + 2  # the second record's vertices start after its header
    with pytest.raises(ValueError, match="finite xyz"):
        s.addHull([0, 0, 0, 1, 0])
    with pytest.raises(ValueError, match="spanning a volume"):
        s.addHull([0, 0, 0, 1, 0, 0, 0, 1, 0])


def test_heightfield_pieces_and_volumes():
    heights = [0.0, 0.2, 0.1, 0.4, 0.3, 0.0, 0.2, 0.1, 0.5]
    cooked = avbd.cookHeightfield(heights=heights, columns=3, rows=3, spacing=[1.0, 2.0], baseHeight=-1.0)
    assert cooked["kind"] == "heightfield" and cooked["triangleCount"] == 8
    H = np.array(heights).reshape(3, 3)
    want = []
    for z in range(2):
        for x in range(2):
            a, b, c, d = H[z, x], H[z, x + 1], H[z + 1, x], H[z + 1, x + 1]
            for tri in ((a, b, c), (b, d, c)):
                want.append(1.0 * 2.0 / 2 * (np.mean(tri) + 1.0))  # area x mean height above the base
    assert np.allclose([p["shape"].volume for p in cooked["pieces"]], want, rtol=1e-9)
    holes = [0, 1, 0, 0]
    assert avbd.cookHeightfield(heights=heights, columns=3, rows=3, holes=holes)["triangleCount"] == 6
    for bad, msg in ((dict(columns=1), "columns"), (dict(heights=heights[:-1]), "heights"),
                     (dict(baseHeight=0.05), "baseHeight"), (dict(upAxis=3), "upAxis"), (dict(holes=[0]), "holes")):
        kw = dict(heights=heights, columns=3, rows=3)
        kw.update(bad)
        with pytest.raises(ValueError, match=msg):
            avbd.cookHeightfield(**kw)


def test_static_mesh_prisms_and_bodies():
    # a ramp of two triangles and a degenerate seam triangle
    pos = [0, 0, 0, 4, 0, 0, 0, 1, 4, 4, 1, 4, 2, 0.5, 2]
    cooked = avbd.cookStaticMesh(positions=pos, indices=[0, 2, 1, 1, 2, 3, 0, 0, 1], thickness=0.2)
    assert cooked["triangleCount"] == 2
    p = np.array(pos).reshape(-1, 3)
    area = 0.5 * np.linalg.norm(np.cross(p[2] - p[0], p[1] - p[0]))
    assert all(abs(pc["shape"].volume - area * 0.2) < 1e-9 for pc in cooked["pieces"])
    s = avbd.AvbdScene3D()
    t = s.addTerrain(cooked, friction=0.8, restitution=0.1)
    assert len(t["bodies"]) == 2 and t["kind"] == "static-mesh"
    for b in t["bodies"]:
        r = s.record(b)
        assert r[19] == 0 and abs(r[3] - 0.8) < 1e-7 and r[39] >= 3
    with pytest.raises(ValueError, match="zero mass"):
        s.addTerrain(cooked, mass=1)
    with pytest.raises(ValueError, match="nondegenerate"):
        avbd.cookStaticMesh(positions=[0, 0, 0, 1, 0, 0, 2, 0, 0])


# --- the oracle's hull contacts ----------------------------------------------------------------------------------------

def _hull_scene(shapes):
    """A scene of hull bodies (shape, position, quaternion) -> (records, hull buffer)."""
    s = avbd.AvbdScene3D(gravity=0)
    for shape, pos, q in shapes:
        if isinstance(shape, str):
            s.addBox([1.0, 0.8, 1.2], position=pos, rotation=q)
        else:
            s.addHullShape(shape, position=pos, rotation=q)
    b, *_ = s.packed()
    return b, s.hull_buffer()


def _random_quat(rng):
    q = rng.normal(size=4)
    return list(q / np.linalg.norm(q))


def _world_geometry(shape, rec):
    """World vertices, face planes (n, d) and edges of a cooked hull body."""
    q = rec[4:8].astype(np.float64)
    c = rec[0:3].astype(np.float64)
    verts = np.array([quat_rotate(q, v) + c for v in shape.vertices.astype(np.float64)])
    planes = []
    for f in shape.faces:
        n = quat_rotate(q, np.array(f["normal"]))
        planes.append((n, f["d"] + n @ c))
    edges = [(verts[a], verts[b]) for a, b, _, _ in shape.edges]
    return verts, planes, edges


def _brute_sat(A, B):
    """Largest separation over every face normal of both and every edge-pair cross product (support functions)."""
    va, pa, ea = A
    vb, pb, eb = B
    axes = [n for n, _ in pa] + [n for n, _ in pb]
    for p0, p1 in ea:
        for q0, q1 in eb:
            n = np.cross(p1 - p0, q1 - q0)
            l = np.linalg.norm(n)
            if l > 1e-9:
                axes.append(n / l)
    best = -np.inf
    for n in axes:
        for s in (n, -n):
            best = max(best, (vb @ s).min() - (va @ s).max())
    return best


SHAPES = {
    "tetra": lambda: avbd.convexHull([-0.5, -0.5, -0.5, 0.5, -0.5, -0.5, 0, 0.5, -0.5, 0, 0, 0.7]),
    "cloud": lambda: avbd.convexHull(np.random.default_rng(5).normal(size=(40, 3)) * [0.6, 0.4, 0.5]),
    "prism": lambda: avbd.hullFromTriangles([0, 0, 0, 1, 0, 0, 0, 0, 1, 0, -0.6, 0, 1, -0.6, 0, 0, -0.6, 1],
                                            avbd.terrain.REVERSED_PRISM_FACES),
}


@pytest.mark.parametrize("pair", [("tetra", "cloud"), ("cloud", "cloud"), ("prism", "tetra"), ("cloud", "prism")])
def test_sat_equals_brute_force_and_points_on_surfaces(pair):
    rng = np.random.default_rng(sum(map(ord, "".join(pair))))
    sa, sb = SHAPES[pair[0]](), SHAPES[pair[1]]()
    touching = separated = 0
    for trial in range(60):
        qa, qb = _random_quat(rng), _random_quat(rng)
        d = rng.normal(size=3)
        d *= rng.uniform(0.2, 2.5) / np.linalg.norm(d)
        recs, buf = _hull_scene([(sa, [0, 0, 0], qa), (sb, list(d), qb)])
        A = _world_geometry(sa, recs[0])
        B = _world_geometry(sb, recs[1])
        brute = _brute_sat(A, B)
        feats, pts, normal, q = probe_hull_pair(recs[0], recs[1], buf, precision=64)
        if brute > 1e-6:
            assert len(feats) == 0, (trial, brute)
            separated += 1
            continue
        if brute > -1e-6:
            continue  # grazing: either answer
        touching += 1
        sat = max(q[0], q[2], q[4] if q[5] else -np.inf)
        # (face planes carry their offsets rounded to f32, which the brute force takes from the vertices)
        assert abs(sat - brute) < 2e-7, (trial, sat, brute)
        assert len(feats) >= 1 and abs(np.linalg.norm(normal) - 1) < 1e-6  # (f32 face normals)
        for xa, xb in zip(pts[:, 0:3], pts[:, 3:6]):
            da = max(n @ xa - w for n, w in A[1])
            db = max(n @ xb - w for n, w in B[1])
            assert abs(da) < 5e-7 and abs(db) < 5e-7, (trial, da, db)
    assert touching >= 10 and separated >= 5, (touching, separated)


def test_hull_of_a_box_collides_like_the_box():
    rng = np.random.default_rng(17)
    shape = avbd.convexHull(box_points(1.0, 0.8, 1.2))
    checked = 0
    for trial in range(80):
        qa, qb = _random_quat(rng), _random_quat(rng)
        d = rng.normal(size=3)
        d *= rng.uniform(0.3, 1.3) / np.linalg.norm(d)
        recs_box, _ = _hull_scene([("box", [0, 0, 0], qa), ("box", list(d), qb)])
        recs_hull, buf = _hull_scene([(shape, [0, 0, 0], qa), (shape, list(d), qb)])
        fb, pb, nb, _ = probe_hull_pair(recs_box[0], recs_box[1], buf, precision=64)
        fh, ph, nh, _ = probe_hull_pair(recs_hull[0], recs_hull[1], buf, precision=64)
        assert (len(fb) == 0) == (len(fh) == 0), trial
        if not len(fb) or (fb[0] >> 24) == 2 or (fh[0] >> 24) == 2:
            continue  # edge contacts pick their edges by different rules
        if abs(abs(nb @ nh) - 1) > 1e-6:
            continue  # nearly equal face separations may pick the other shape's face
        checked += 1
        assert np.allclose(nb, nh, atol=1e-6)
        # the same points (the hull's face data are f32-rounded: within a few 1e-6)
        assert len(pb) == len(ph), trial
        for p in pb:
            assert np.min(np.linalg.norm(ph - p, axis=1)) < 3e-6, trial
    assert checked >= 15, checked


def _polytope_distance(planes, point):
    """Distance from a point to a convex polytope {x: n.x <= d} (0 inside), by constrained minimisation."""
    cons = [{"type": "ineq", "fun": (lambda x, n=n, w=w: w - n @ x)} for n, w in planes]
    r = minimize(lambda x: np.sum((x - point) ** 2), point * 0.0, constraints=cons, method="SLSQP",
                 options=dict(ftol=1e-14, maxiter=500))
    return math.sqrt(max(r.fun, 0.0)), r.x


def test_sphere_contact_at_the_polytope_distance():
    rng = np.random.default_rng(23)
    shape = SHAPES["cloud"]()
    hits = 0
    for trial in range(40):
        q = _random_quat(rng)
        d = rng.normal(size=3)
        d *= rng.uniform(0.3, 1.2) / np.linalg.norm(d)
        s = avbd.AvbdScene3D(gravity=0)
        s.addHullShape(shape, position=[0, 0, 0], rotation=q)
        s.addSphere(0.35, position=list(d))
        recs, *_ = s.packed()
        buf = s.hull_buffer()
        _, planes, _ = _world_geometry(shape, recs[0])
        dist, closest = _polytope_distance(planes, recs[1][0:3].astype(np.float64))
        feats, pts, normal, _ = probe_hull_pair(recs[0], recs[1], buf, precision=64)
        if dist > 0.35 + 1e-6:
            assert len(feats) == 0
            continue
        if dist < 1e-6 or dist > 0.35 - 1e-6:
            continue  # centre inside (the shallowest face rule), or grazing
        hits += 1
        on_hull = pts[0, 0:3]  # A is the hull
        assert np.linalg.norm(on_hull - closest) < 1e-5, (trial, on_hull, closest)
    assert hits >= 8, hits


def _segment_polytope_distance(planes, a, b):
    cons = [{"type": "ineq", "fun": (lambda z, n=n, w=w: w - n @ z[:3])} for n, w in planes]
    cons.append({"type": "ineq", "fun": lambda z: z[3]})
    cons.append({"type": "ineq", "fun": lambda z: 1 - z[3]})
    best = None
    for t0 in (0.0, 0.5, 1.0):
        r = minimize(lambda z: np.sum((z[:3] - (a + z[3] * (b - a))) ** 2), np.r_[np.zeros(3), t0],
                     constraints=cons, method="SLSQP", options=dict(ftol=1e-15, maxiter=500))
        if best is None or r.fun < best:
            best = r.fun
    return math.sqrt(max(best, 0.0))


def test_capsule_contact_at_the_segment_distance():
    rng = np.random.default_rng(29)
    shape = SHAPES["tetra"]()
    hits = 0
    for trial in range(40):
        q = _random_quat(rng)
        d = rng.normal(size=3)
        d *= rng.uniform(0.4, 1.1) / np.linalg.norm(d)
        s = avbd.AvbdScene3D(gravity=0)
        s.addCapsule(0.2, 1.0, position=list(d), rotation=_random_quat(rng))
        s.addHullShape(shape, position=[0, 0, 0], rotation=q)
        recs, *_ = s.packed()
        buf = s.hull_buffer()
        _, planes, _ = _world_geometry(shape, recs[1])
        c = recs[0]
        axis = quat_rotate(c[4:8].astype(np.float64), np.array([0, 0.5 - 0.2, 0]))
        a, b = c[0:3] - axis, c[0:3] + axis
        dist = _segment_polytope_distance(planes, a, b)
        feats, pts, normal, _ = probe_hull_pair(recs[0], recs[1], buf, precision=64)
        if dist > 0.2 + 1e-5:
            assert len(feats) == 0, (trial, dist)
            continue
        if dist < 1e-4 or dist > 0.2 - 1e-5:
            continue
        hits += 1
        # xA = segment point + n r, xB = the hull point: their gap along n is the distance less the radius
        sep = (pts[0, 3:6] - pts[0, 0:3]) @ (-normal)
        assert abs(sep - (dist - 0.2)) < 1e-5, (trial, sep, dist)
    assert hits >= 6, hits


# --- the oracle with hulls: rest heights --------------------------------------------------------------------------------

def test_hulls_and_terrain_rest_in_the_oracle():
    s = avbd.AvbdScene3D(gravity=-9.81)
    s.addHeightfield(heights=[0.0] * 16, columns=4, rows=4, spacing=2.0, origin=[-3, 0, -3])
    cube = s.addHull(box_points(1, 1, 1), position=[0.2, 0.8, 0.1])
    ball = s.addSphere(0.3, position=[-1.5, 0.6, -1.2])
    caps = s.addCapsule(0.25, 1.2, position=[1.5, 0.6, 1.4], rotation=[0, 0, math.sin(math.pi / 4), math.cos(math.pi / 4)])
    box = s.addBox([0.6, 0.6, 0.6], position=[-1.4, 0.6, 1.5])
    o = OracleWorld([s], precision=32)
    o.step(150)
    b = o.bodies()
    # each rests about one 0.01 m margin deep on the flat terrain (top at 0)
    assert abs(b[cube, 1] - 0.49) < 0.01, b[cube, 1]
    assert abs(b[ball, 1] - 0.29) < 0.01, b[ball, 1]
    # a capsule gets one point per hull (capsulePoly): it settles lower, within the reference's own bound for a capsule
    # on a hull floor, 0.06 m (tests/gpu-package-features.js:262-268)
    assert 0.25 - 0.06 < b[caps, 1] < 0.25, b[caps, 1]
    assert abs(b[box, 1] - 0.29) < 0.01, b[box, 1]
    assert np.abs(b[[cube, ball, caps, box], 32:35]).max() < 0.05


# --- shape queries against hulls (the oracle's roundedPoly for hull records) ---------------------------------------------

def _cast_scene(build):
    from oracle_lib import cast
    s = avbd.AvbdScene3D(gravity=0)
    build(s)
    b, *_ = s.packed()
    filters = np.full((len(b), 2), 0xFFFFFFFF, np.uint32)
    return lambda q: cast(b, filters, np.asarray(q, np.float32).reshape(-1, 8), hulls=s.hull_buffer())


def test_sphere_casts_round_a_hull_corner():
    # tests/gpu-package-features.js:193-221: a 2 m hull cube, a 0.5 sphere swept along +x
    run = _cast_scene(lambda s: s.addHull(box_points(2, 2, 2), density=0))
    hit = run([-3, 0, 0, 1e3, 1, 0, 0, 0.5])[0]
    assert abs(hit[3] - 1.5) < 1e-3 and np.allclose(hit[0:3], [-1, 0, 0])
    assert run([-3, 1.4, 1.4, 1e3, 1, 0, 0, 0.5])[0][3] < 0, "the hull's corner is rounded, not its bounding box"


@pytest.mark.parametrize("radius", [0.0, 0.3])
def test_queries_of_a_box_hull_equal_the_box(radius):
    rng = np.random.default_rng(41)
    hull_run = _cast_scene(lambda s: s.addHull(box_points(1.4, 0.8, 2.0), density=0))
    box_run = _cast_scene(lambda s: s.addBox([1.4, 0.8, 2.0], density=0))
    o = rng.uniform(-3, 3, (400, 3))
    d = rng.uniform(-1, 1, (400, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    q = np.concatenate([o, np.full((400, 1), 1e3), d, np.full((400, 1), radius)], 1)
    h, b = hull_run(q), box_run(q)
    assert np.array_equal(h[:, 3] >= 0, b[:, 3] >= 0)
    hit = b[:, 3] >= 0
    assert 100 < hit.sum() < 400
    # (an edge's swept capsule is solved from its other end when the hull lists it the other way round: the FP32
    # quadratic then rounds differently, by up to a few 1e-6 in distance and 1e-5 in normal)
    assert np.allclose(h[hit, 3], b[hit, 3], atol=1e-5)
    assert np.allclose(h[hit, 0:3], b[hit, 0:3], atol=3e-5)


def test_rays_into_a_hull_match_a_slab_clip():
    rng = np.random.default_rng(43)
    shape = avbd.convexHull(rng.normal(size=(25, 3)) * [0.8, 0.5, 0.6])
    q0 = _random_quat(rng)
    run = _cast_scene(lambda s: s.addHullShape(shape, density=0, position=[0.3, -0.2, 0.1], rotation=q0))
    s = avbd.AvbdScene3D()
    s.addHullShape(shape, density=0, position=[0.3, -0.2, 0.1], rotation=q0)
    rec = s.packed()[0][0]
    _, planes, _ = _world_geometry(shape, rec)
    o = rng.uniform(-3, 3, (300, 3))
    d = rng.normal(size=(300, 3)) * 0.3 - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    hits = run(np.concatenate([o, np.full((300, 1), 1e3), d, np.zeros((300, 1))], 1))
    agree = 0
    for k in range(300):
        enter, leave = -np.inf, np.inf
        for n, w in planes:  # the ray's entry and exit through every face plane (Cyrus–Beck), in float64
            dn, gap = n @ d[k], w - n @ o[k]
            if abs(dn) < 1e-12:
                if gap < 0:
                    enter, leave = 1, 0
            elif dn < 0:
                enter = max(enter, gap / dn)
            else:
                leave = min(leave, gap / dn)
        want = None if enter > leave or leave < 0 else max(enter, 0.0)
        got = hits[k, 3]
        if want is None:
            assert got < 0 or leave - enter > -1e-5, k
        else:
            assert got >= 0 and abs(got - want) < 1e-5, (k, got, want)
            agree += 1
    assert agree > 50

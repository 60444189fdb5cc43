"""Pins of the GPU-semantics oracle (oracle/gpusem.hpp) by checks independent of its own formulas (CPU only):

* box contacts vs an independent convex distance (a QP over the two boxes), surface membership of contact points, and
  the 8-vertex clipping bound;
* the pair set: grid broadphase == brute force, and every touching pair present;
* adjacency == an independent construction; colouring proper, and on a fresh graph equal to sequential greedy in
  descending lowbias32 priority (what Jones–Plassmann reduces to);
* the 6x6 LDLᵀ against numpy;
* the reference's physical known answers (avbd3d-ref.test.ts:51-102, gpu-canonical.js:39-65) run through the GPU
  semantics: resting height 1 − margin with ≥ 4 contacts, stack rest, Coulomb stopping distance, spring equilibrium,
  rope joint error.
"""
import ctypes
import math

import numpy as np
import pytest
from scipy.optimize import minimize

from oracle_lib import OracleWorld, lib

D = ctypes.POINTER(ctypes.c_double)
U = ctypes.POINTER(ctypes.c_uint32)


def quat_to_mat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def random_quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def box_contacts(pa, qa, sa, pb, qb, sb, face_bias=1, max_poly=8, precision=64):
    poseA = np.array([*pa, *qa], np.float64)
    poseB = np.array([*pb, *qb], np.float64)
    sizeA = np.asarray(sa, np.float64)
    sizeB = np.asarray(sb, np.float64)
    feats = np.zeros(8, np.uint32)
    pts = np.zeros((8, 6))
    normal = np.zeros(3)
    n = lib().gs_probe_box_contacts(precision, poseA.ctypes.data_as(D), sizeA.ctypes.data_as(D),
                                    poseB.ctypes.data_as(D), sizeB.ctypes.data_as(D), face_bias, max_poly,
                                    feats.ctypes.data_as(U), pts.ctypes.data_as(D), normal.ctypes.data_as(D))
    return feats[:n], pts[:n], normal


def box_distance(pa, qa, sa, pb, qb, sb):
    """Distance between two boxes by a QP over local coordinates (0 when they overlap)."""
    Ra, Rb = quat_to_mat(qa), quat_to_mat(qb)
    ha, hb = np.asarray(sa) / 2, np.asarray(sb) / 2

    def f(x):
        d = (pa + Ra @ x[:3]) - (pb + Rb @ x[3:])
        return d @ d

    def g(x):
        d = (pa + Ra @ x[:3]) - (pb + Rb @ x[3:])
        return np.concatenate([2 * Ra.T @ d, -2 * Rb.T @ d])

    bounds = [(-h, h) for h in ha] + [(-h, h) for h in hb]
    best = None
    for start in (np.zeros(6), np.concatenate([Ra.T @ (pb - pa), Rb.T @ (pa - pb)])):
        start = np.clip(start, [b[0] for b in bounds], [b[1] for b in bounds])
        r = minimize(f, start, jac=g, bounds=bounds, method="L-BFGS-B", options=dict(ftol=1e-15, gtol=1e-12))
        if best is None or r.fun < best.fun:
            best = r
    return math.sqrt(max(best.fun, 0.0))


def random_box_pair(rng, gap_scale=0.6):
    sa = rng.uniform(0.3, 2.0, 3)
    sb = rng.uniform(0.3, 2.0, 3)
    qa, qb = random_quat(rng), random_quat(rng)
    pa = rng.uniform(-1, 1, 3)
    direction = rng.normal(size=3)
    direction /= np.linalg.norm(direction)
    reach = (np.linalg.norm(sa) + np.linalg.norm(sb)) / 2
    pb = pa + direction * reach * rng.uniform(0.2, gap_scale * 2)
    return pa, qa, sa, pb, qb, sb


def test_sat_agrees_with_an_independent_box_distance():
    rng = np.random.default_rng(1)
    checked = 0
    for _ in range(400):
        pa, qa, sa, pb, qb, sb = random_box_pair(rng)
        dist = box_distance(pa, qa, sa, pb, qb, sb)
        feats, pts, _ = box_contacts(pa, qa, sa, pb, qb, sb)
        if dist > 1e-4:
            assert len(feats) == 0, f"separated by {dist} but {len(feats)} contacts"
            checked += 1
        elif dist == 0.0:
            # overlapping (QP reaches zero): SAT must find no separating axis
            pass
    assert checked > 50


def test_contact_points_lie_on_the_box_surfaces():
    rng = np.random.default_rng(2)
    seen = 0
    for _ in range(300):
        pa, qa, sa, pb, qb, sb = random_box_pair(rng, gap_scale=0.45)
        feats, pts, normal = box_contacts(pa, qa, sa, pb, qb, sb)
        if not len(feats):
            continue
        seen += 1
        assert abs(np.linalg.norm(normal) - 1) < 1e-9
        for (xa, xb) in zip(pts[:, :3], pts[:, 3:]):
            for x, p, q, s in ((xa, pa, qa, sa), (xb, pb, qb, sb)):
                u = quat_to_mat(q).T @ (x - p)
                excess = np.abs(u) - np.asarray(s) / 2
                assert excess.max() < 1e-6, "contact point outside its box"
                assert excess.max() > -1e-5, "contact point not on its box's surface"
        kinds = feats >> 24
        assert set(kinds.tolist()) <= {0, 1, 2}
    assert seen > 50


def test_eight_clipped_vertices_suffice():
    """A quad clipped by four planes has at most 8 vertices: the GPU's 8-slot clip equals the demo's 16-slot one."""
    rng = np.random.default_rng(3)
    for _ in range(300):
        args = random_box_pair(rng, gap_scale=0.4)
        f8, p8, n8 = box_contacts(*args, max_poly=8)
        f16, p16, n16 = box_contacts(*args, max_poly=16)
        assert np.array_equal(f8, f16) and np.array_equal(p8, p16) and np.array_equal(n8, n16)


def test_face_bias_only_changes_the_axis_choice():
    """With the face bias off the narrowphase is the demo's collide (pinned by the golden trajectories); on, every
    result is still a valid contact set (points on surfaces) and face manifolds are preferred."""
    rng = np.random.default_rng(4)
    edges_off = edges_on = 0
    for _ in range(300):
        args = random_box_pair(rng, gap_scale=0.3)
        f0, _, _ = box_contacts(*args, face_bias=0)
        f1, _, _ = box_contacts(*args, face_bias=1)
        edges_off += int(((f0 >> 24) == 2).any())
        edges_on += int(((f1 >> 24) == 2).any())
    assert edges_on <= edges_off


def _pile(avbd, n=(13, 13, 13), seed=5):
    return avbd.configs.box_pile(counts=n, seed=seed, centre=(0, 8.0, 0))


def test_grid_pairs_equal_brute_force(avbd):
    grid = OracleWorld([_pile(avbd)], precision=32)
    brute = OracleWorld([_pile(avbd)], precision=32, brute_force=True)
    for _ in range(3):
        grid.step(20)
        brute.step(20)
        pg = {tuple(p) for p in grid.pairs().tolist()}
        pb = {tuple(p) for p in brute.pairs().tolist()}
        assert pg == pb and len(pg) > 1000


def test_every_touching_pair_is_found(avbd):
    """Bodies whose boxes are within 1e-4 (independent QP distance) are always in the pair set."""
    w = OracleWorld([avbd.configs.box_pile(counts=(5, 5, 5), seed=9)], precision=64)
    w.step(60)
    b = w.bodies()
    pairs = {tuple(p) for p in w.pairs().tolist()}
    n = len(b)
    checked = 0
    for a in range(1, n):
        for c in range(1, a):
            d = b[a, 0:3] - b[c, 0:3]
            if d @ d > (b[a, 23] + b[c, 23]) ** 2:
                continue
            dist = box_distance(b[a, 0:3], b[a, 4:8], b[a, 16:19], b[c, 0:3], b[c, 4:8], b[c, 16:19])
            if dist < 1e-4:
                assert (a, c) in pairs, f"touching pair {(a, c)} missing"
                checked += 1
    assert checked > 50


def lowbias32(x):
    x &= 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def test_priority_is_lowbias32():
    for i in [0, 1, 2, 77, 1000, 123456]:
        assert lib().gs_probe_priority(i) == lowbias32(i + 0x9E3779B9)


def _graph(w):
    """Neighbour sets of dynamic bodies from the oracle's manifolds (independent of its adjacency)."""
    b = w.bodies()
    dyn = b[:, 19] > 0
    nbr = {i: set() for i in range(len(b)) if dyn[i]}
    for row in w.manifolds():
        a, c = int(row[0]), int(row[1])
        if dyn[a] and dyn[c]:
            nbr[a].add(c)
            nbr[c].add(a)
    return nbr, dyn


def test_adjacency_is_sorted_and_complete(avbd):
    w = OracleWorld([avbd.configs.box_pile(counts=(6, 6, 6), seed=3)], precision=32)
    w.step(40)
    start, lst = w.adjacency()
    mf = w.manifolds()
    b = w.bodies()
    expect = {i: [] for i in range(len(b))}
    for m, row in enumerate(mf):
        for x, other in ((int(row[0]), int(row[1])), (int(row[1]), int(row[0]))):
            if b[x, 19] > 0:
                expect[x].append((other + 2, m))
    for i in range(len(b)):
        got = lst[start[i]:start[i + 1]].tolist()
        want = [m for _, m in sorted(expect[i])]
        assert got == want, f"body {i}"


def test_first_colouring_is_greedy_in_priority_order(avbd):
    w = OracleWorld([avbd.configs.box_pile(counts=(8, 8, 8), seed=11)], precision=32, color_cap=64, color_rounds=32)
    w.step(1)
    assert w.counts()["clashes"] == 0
    nbr, dyn = _graph(w)
    order = sorted(nbr, key=lambda i: (lowbias32(i + 0x9E3779B9), i), reverse=True)
    greedy = {}
    for i in order:
        used = {greedy[j] for j in nbr[i] if j in greedy}
        c = 0
        while c in used:
            c += 1
        greedy[i] = c
    colors = w.colors()
    for i in nbr:
        assert colors[i] == greedy[i], f"body {i}: {colors[i]} vs greedy {greedy[i]}"


def test_colouring_stays_proper(avbd):
    w = OracleWorld([avbd.configs.box_pile(counts=(8, 8, 8), seed=12)], precision=32, color_cap=32, color_rounds=16)
    for _ in range(6):
        w.step(15)
        assert w.counts()["clashes"] == 0
        nbr, _ = _graph(w)
        colors = w.colors()
        for i, ns in nbr.items():
            assert all(colors[i] != colors[j] for j in ns)


def test_ldlt_matches_numpy():
    rng = np.random.default_rng(7)
    for _ in range(200):
        M = rng.normal(size=(6, 6))
        A = M @ M.T + 0.1 * np.eye(6)
        b = rng.normal(size=6)
        x = np.zeros(6)
        lib().gs_probe_solve6(A.ravel().ctypes.data_as(D), b.ctypes.data_as(D), 0.0, x.ctypes.data_as(D))
        want = np.linalg.solve(A, b)
        assert np.abs(x - want).max() <= 1e-9 * max(1.0, np.abs(want).max())
    # The pivot floor keeps a singular direction finite
    A = np.eye(6)
    A[1, 1] = A[0, 0]
    A[1, 0] = A[0, 1] = 1.0  # rank-deficient 2x2 block
    x = np.zeros(6)
    lib().gs_probe_solve6(A.ravel().ctypes.data_as(D), np.ones(6).ctypes.data_as(D), 1e-5, x.ctypes.data_as(D))
    assert np.isfinite(x).all()


# --- the reference's physical known answers through the GPU semantics (z up, g = -10) ----------------------------


def zscene(avbd, **kw):
    return avbd.AvbdScene3D(gravity=-10.0, up=(0, 0, 1), **kw)


def test_resting_contact_respects_the_margin(avbd):
    """gpu-canonical.js:39-65."""
    s = zscene(avbd)
    s.addBox([100, 100, 1], density=0, friction=0.5, position=[0, 0, 0])
    s.addBox([1, 1, 1], friction=0.5, position=[0, 0, 4])
    w = OracleWorld([s], precision=32)
    w.step(240)
    z = w.bodies()[1, 2]
    assert abs(z - (1 - 0.01)) < 0.01
    c = w.counts()
    assert c["contacts"] >= 4 and c["clashes"] == 0


def test_stack_rests_within_the_margin(avbd):
    s = zscene(avbd)
    s.addBox([100, 100, 1], density=0, friction=0.5, position=[0, 0, 0])
    for i in range(10):
        s.addBox([1, 1, 1], friction=0.5, position=[0, 0, i * 1.5 + 1])
    w = OracleWorld([s], precision=64)
    w.step(900)
    b = w.bodies()
    for i in range(10):
        assert abs(b[1 + i, 2] - (i + 1)) < 0.011 * (i + 1) + 0.01, f"box {i} z {b[1 + i, 2]}"
        assert np.linalg.norm(b[1 + i, 32:35]) < 1e-2


def test_dynamic_friction_follows_coulomb(avbd):
    s = zscene(avbd)
    s.addBox([100, 100, 1], density=0, friction=0.5, position=[0, 0, 0])
    for x in range(11):
        s.addBox([1, 1, 0.5], friction=5 - (x / 10) * 5, position=[0, -30 + x * 2, 0.75], velocity=[10, 0, 0])
    w = OracleWorld([s], precision=64)
    w.step(300)
    b = w.bodies()
    for i in range(10):
        mu = math.sqrt((5 - i * 0.5) * 0.5)
        expected = 100 / (2 * mu * 10)
        assert abs(b[1 + i, 0] - expected) / expected < 0.15, f"box {i}: slid {b[1 + i, 0]}, Coulomb {expected}"
    assert abs(b[11, 0] - 50) < 1e-3


def test_spring_oscillates_about_its_static_equilibrium(avbd):
    s = zscene(avbd)
    s.addBox([100, 100, 1], density=0, friction=0.5, position=[0, 0, 0])
    anchor = s.addBox([1, 1, 1], density=0, friction=0.5, position=[0, 0, 14])
    block = s.addBox([2, 2, 2], friction=0.5, position=[0, 0, 8])
    s.addSpring(anchor, block, stiffness=100, rest=4)
    w = OracleWorld([s], precision=64)
    w.step(600)
    total = 0.0
    for _ in range(600):
        w.step(1)
        total += w.bodies()[2, 2]
    assert abs(total / 600 - 9.2) < 0.05


def test_rope_joints_hold(avbd):
    s = zscene(avbd)
    s.addBox([100, 100, 1], density=0, friction=0.5, position=[0, 0, -20])
    prev = None
    for i in range(20):
        cur = s.addBox([1, 0.5, 0.5], density=0 if i == 0 else 1, friction=0.5, position=[i, 0, 10])
        if prev is not None:
            s.addJoint(prev, cur, [0.5, 0, 0], [-0.5, 0, 0])
        prev = cur
    w = OracleWorld([s], precision=64)
    w.step(240)
    b = w.bodies()

    def rot(q, v):
        return quat_to_mat(q) @ np.asarray(v)
    err = 0.0
    for i in range(1, 20):
        pa = b[i, 0:3] + rot(b[i, 4:8], [0.5, 0, 0])
        pb = b[i + 1, 0:3] + rot(b[i + 1, 4:8], [-0.5, 0, 0])
        err = max(err, np.linalg.norm(pa - pb))
    assert err < 0.02, f"joint error {err}"


@pytest.mark.parametrize("cap", [16, 3])
def test_colour_parallel_variant_equals_in_order(avbd, cap):
    """The OpenMP variant of the oracle (the all-cores CPU timing) steps to the same bytes as one thread: bodies of a
    colour, pairs and manifolds are independent, and a colouring with clashes (cap 3 forces them) solves in order."""
    from oracle_lib import get_threads, set_threads
    scenes = [avbd.configs.pile_of(400, seed=11), avbd.configs.jointed_mix(n=8, spheres=60, capsules=60, area=8.0)]
    out = []
    for threads in (1, 4):
        set_threads(threads)
        try:
            assert get_threads() == threads
            w = OracleWorld(scenes, precision=32, color_cap=cap, color_rounds=4)
            w.step(40)
            out.append((w.bodies(), w.joints(), w.contact_words(), w.counts()))
        finally:
            set_threads(1)
    (b1, j1, c1, n1), (b4, j4, c4, n4) = out
    assert n1 == n4
    if cap == 3:
        assert n1["clashes"] > 0
    np.testing.assert_array_equal(b1, b4)
    np.testing.assert_array_equal(j1, j4)
    np.testing.assert_array_equal(c1, c4)


def _low_pile():
    import importlib
    avbd = importlib.import_module("babylon-avbd_amd")
    return avbd.configs.box_pile(counts=(4, 2, 4), centre=(0.0, 1.2, 0.0), seed=5)


@pytest.mark.parametrize("failure", ["clashes", "overflow"])
def test_failed_step_is_held(failure):
    """protectSolverStep (src/gpu/gpuSolverKernels.js:72-103): a step whose colouring clashes (one colour for touching
    bodies) or whose pairs exceed the pair capacity leaves every body as it was; the same step with room solves."""
    free = OracleWorld([_low_pile()], precision=32, color_cap=16, color_rounds=16)
    tight = OracleWorld([_low_pile()], precision=32, color_cap=1 if failure == "clashes" else 16, color_rounds=2)
    if failure == "overflow":
        tight.set_pair_capacity(3)
    for w in (free, tight):
        w.step(6)  # the pile lands: contacts between the boxes and with the ground
    before = tight.bodies().copy()
    free.seed(dict(bodies=before.astype(np.float32), joints=tight.joints().astype(np.float32), refs=tight.refs(),
                   colors=tight.colors(), manifolds=tight.manifolds(), contacts=tight.contact_words(),
                   steps=np.array([6], np.uint32)))
    tight.step()
    free.step()
    c = tight.counts()
    assert c[failure] > 0
    assert free.counts()["clashes"] == 0 and free.counts()["overflow"] == 0
    held = tight.bodies()
    assert np.array_equal(held[:, np.r_[0:8, 32:40]], before[:, np.r_[0:8, 32:40]])
    assert not np.array_equal(free.bodies()[:, 0:8], before[:, 0:8])

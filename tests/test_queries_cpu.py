"""The oracle's shape queries (src/gpu/shapeQueries.wgsl, shapeQueries.js), pinned independently of its formulas:
a signed-distance ray march (the textbook distance functions of a box, a sphere and a capsule, evaluated in float64
along the ray) must find the same first contact distance and a normal along the distance function's gradient; closed
forms for axis-aligned cases; the filters, the ignore list and the smallest-index rule on ties."""
import numpy as np
import pytest

from oracle_lib import cast


def body(shape, pos, rot, size):
    r = np.zeros(40, np.float32)
    r[0:3], r[4:8], r[16:19] = pos, rot, size
    r[19] = 1.0
    r[39] = {"box": 0.0, "sphere": 1.0, "capsule": -1.0}[shape]
    return r


def rotation(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def sdf(rec, p):
    """Distance from world points p (..., 3) to the body's surface (negative inside), float64."""
    R = rotation(rec[4:8].astype(np.float64))
    local = (np.asarray(p, np.float64) - rec[0:3].astype(np.float64)) @ R
    size = rec[16:19].astype(np.float64)
    if rec[39] == 1.0:
        return np.linalg.norm(local, axis=-1) - size[0] / 2
    if rec[39] == -1.0:
        half = max(0.0, size[1] / 2 - size[0] / 2)
        c = np.zeros_like(local)
        c[..., 1] = np.clip(local[..., 1], -half, half)
        return np.linalg.norm(local - c, axis=-1) - size[0] / 2
    q = np.abs(local) - size / 2
    return np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(axis=-1), 0.0)


def march(rec, o, d, r, tmax=12.0, step=1e-3):
    """First t with sdf(o + t d) <= r by a fine march refined by bisection, and the smallest clearance seen."""
    ts = np.arange(0, tmax, step)
    vals = sdf(rec, o[None, :] + ts[:, None] * d[None, :]) - r
    inside = np.nonzero(vals <= 0)[0]
    if not len(inside):
        return None, vals.min()
    k = inside[0]
    if k == 0:
        return 0.0, vals.min()
    lo, hi = ts[k - 1], ts[k]
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if sdf(rec, o + mid * d) - r <= 0:
            hi = mid
        else:
            lo = mid
    return hi, vals.min()


def unit_quat(rng):
    q = rng.normal(size=4)
    return (q / np.linalg.norm(q)).astype(np.float32)


@pytest.mark.parametrize("shape", ["box", "sphere", "capsule"])
@pytest.mark.parametrize("radius", [0.0, 0.15, 0.4])
def test_casts_agree_with_a_signed_distance_march(shape, radius):
    rng = np.random.default_rng(100 * ["box", "sphere", "capsule"].index(shape) + int(radius * 100))
    checked = 0
    for trial in range(40):
        size = rng.uniform(0.4, 1.6, 3).astype(np.float32)
        if shape == "capsule":
            size[1] = size[0] + rng.uniform(0.2, 1.5)
        rec = body(shape, rng.uniform(-1, 1, 3), unit_quat(rng), size)
        o = rng.normal(size=3)
        o = o / np.linalg.norm(o) * rng.uniform(3.0, 5.0) + rec[0:3]
        target = rec[0:3] + rng.uniform(-0.8, 0.8, 3)
        d = (target - o) / np.linalg.norm(target - o)
        q = np.concatenate([o, [100.0], d, [radius]]).astype(np.float32)
        hit = cast(rec[None], np.array([[0xFFFFFFFF, 0xFFFFFFFF]]), q[None])[0]
        t_ref, clearance = march(rec, q[0:3].astype(np.float64), q[4:7].astype(np.float64), radius)
        if abs(clearance) < 5e-3:  # grazing: hit or miss is decided by rounding
            continue
        checked += 1
        if t_ref is None:
            assert hit[3] == -1, f"trial {trial}: oracle hit at {hit[3]}, the march found none"
            continue
        assert abs(hit[3] - t_ref) < 2e-4 * max(1.0, t_ref), f"trial {trial}: {hit[3]} vs {t_ref}"
        # the normal points along the distance function's gradient at the swept sphere's centre
        c = q[0:3].astype(np.float64) + hit[3] * q[4:7].astype(np.float64)
        g = np.array([sdf(rec, c + e) - sdf(rec, c - e) for e in np.eye(3) * 1e-5]) / 2e-5
        g /= np.linalg.norm(g)
        assert np.dot(g, hit[0:3]) > 0.995, f"trial {trial}: normal {hit[0:3]} vs gradient {g}"
    assert checked >= 30


def test_closed_forms():
    box = body("box", [0, 0, 0], [0, 0, 0, 1], [1, 1, 1])
    ball = body("sphere", [0, 0, 0], [0, 0, 0, 1], [1, 1, 1])
    f = np.array([[0xFFFFFFFF, 0xFFFFFFFF]])
    down = lambda r: np.array([[0, 5, 0, 100, 0, -1, 0, r]], np.float32)
    for rec, r, t in [(box, 0.0, 4.5), (box, 0.25, 4.25), (ball, 0.0, 4.5), (ball, 0.5, 4.0)]:
        h = cast(rec[None], f, down(r))[0]
        assert h[3] == np.float32(t) and np.allclose(h[0:3], [0, 1, 0]), (r, h)
    # starting inside: distance 0, normal -direction; beyond the maximum distance: no hit
    h = cast(box[None], f, np.array([[0.1, 0.2, 0, 100, 1, 0, 0, 0]], np.float32))[0]
    assert h[3] == 0 and np.allclose(h[0:3], [-1, 0, 0])
    h = cast(box[None], f, np.array([[0, 5, 0, 4.4, 0, -1, 0, 0]], np.float32))[0]
    assert h[3] == -1 and h[4:5].view(np.uint32)[0] == 0xFFFFFFFF


def test_filters_ignore_and_ties():
    a = body("sphere", [0, 0, 0], [0, 0, 0, 1], [1, 1, 1])
    near = body("box", [0, 2, 0], [0, 0, 0, 1], [1, 1, 1])
    recs = np.stack([a, a.copy(), near])
    f = np.array([[1, 0xFFFFFFFF], [1, 0xFFFFFFFF], [2, 4]], np.uint32)
    q = np.array([[0, 5, 0, 100, 0, -1, 0, 0]], np.float32)
    idx = lambda h: int(h[0, 4:5].view(np.uint32)[0])
    assert idx(cast(recs, f, q)) == 2                               # the nearest body
    assert idx(cast(recs, f, q, mask=1)) == 0                       # group 2 masked out; the tie keeps index 0
    assert idx(cast(recs, f, q, mask=1, ignore=[0])) == 1
    assert idx(cast(recs, f, q, membership=4)) == 2                 # only body 2's mask takes membership 4
    assert cast(recs, f, q, membership=0)[0, 3] == -1               # membership 0 sees nothing

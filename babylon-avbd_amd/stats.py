"""Long-horizon statistics of a world's state: rest heights, energy and penetration (numpy, from the records the world
API reads back: bodies, manifolds and contacts).

* rest height: each dynamic body's centre height along `up`, and its lowest point (boxes: the lowest of the eight
  corners; spheres and capsules: centre or segment end minus radius), summarised as mean / min / percentiles;
* energy: kinetic ½ m |v|² + ½ ωᵀ I_w ω with I_w = R diag(I) Rᵀ, and potential -m g h (g the signed acceleration along
  `up`, so a fall lowers it), summed over dynamic bodies;
* penetration: per contact point, the separation of its two anchors along the pair normal at the current poses,
  dot(n, (x_A + R_A r_A) - (x_B + R_B r_B)) (the normal points from B to A, as the contact rows' C0 = that
  separation + the 0.01 m margin, reference/three-avbd/src/avbd3d/gpu/wgsl-collision.ts:741-807), reported as depth = -separation: max,
  mean and the 99th percentile; a solved contact rests near one margin deep.
"""
import numpy as np

SHAPE_SPHERE, SHAPE_CAPSULE = 1.0, -1.0


def _rotate(q, v):
    """Rotate rows of v by unit quaternions q (x, y, z, w)."""
    u = q[:, 0:3]
    w = q[:, 3:4]
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def lowest_points(bodies, up):
    """Height along `up` of each body's lowest point."""
    b = np.asarray(bodies, np.float64)
    up = np.asarray(up, np.float64)
    centre = b[:, 0:3] @ up
    q = b[:, 4:8]
    half = b[:, 16:19] * 0.5
    shape = b[:, 39]
    # boxes: the box's extent along up is |R^T up| . half
    local_up = _rotate(q * np.array([-1, -1, -1, 1.0]), np.broadcast_to(up, (len(b), 3)))
    box_low = centre - (np.abs(local_up) * half).sum(1)
    radius = half[:, 0]
    sphere_low = centre - radius
    seg = np.maximum(half[:, 1] - radius, 0.0)
    axis_up = _rotate(q, np.broadcast_to(np.array([0, 1.0, 0]), (len(b), 3))) @ up
    capsule_low = centre - np.abs(axis_up) * seg - radius
    return np.where(shape == SHAPE_SPHERE, sphere_low, np.where(shape == SHAPE_CAPSULE, capsule_low, box_low))


def energy(bodies, up, gravity):
    """(kinetic, potential) summed over dynamic bodies; potential relative to height 0."""
    b = np.asarray(bodies, np.float64)
    dyn = b[:, 19] > 0
    b = b[dyn]
    m = b[:, 19]
    v = b[:, 32:35]
    w = b[:, 36:39]
    q = b[:, 4:8]
    w_local = _rotate(q * np.array([-1, -1, -1, 1.0]), w)
    kinetic = 0.5 * (m * (v * v).sum(1)).sum() + 0.5 * ((b[:, 20:23] * w_local * w_local).sum(1)).sum()
    h = b[:, 0:3] @ np.asarray(up, np.float64)
    potential = float((-m * gravity * h).sum())
    return float(kinetic), potential


def contact_depths(bodies, manifolds, contacts):
    """Depth (-separation) of every contact point at the current poses."""
    b = np.asarray(bodies, np.float64)
    mf = np.asarray(manifolds, np.uint32).reshape(-1, 8)
    ct = np.asarray(contacts, np.uint32).reshape(-1, 16)
    if not len(mf):
        return np.zeros(0)
    counts = (mf[:, 3] & 15).astype(np.int64)
    firsts = mf[:, 2].astype(np.int64)
    within = np.arange(int(counts.sum())) - np.repeat(np.cumsum(counts) - counts, counts)
    idx = np.repeat(firsts, counts) + within
    a = np.repeat(mf[:, 0].astype(np.int64), counts)
    bb = np.repeat(mf[:, 1].astype(np.int64), counts)
    normal = np.repeat(mf[:, 4:7].view(np.float32).astype(np.float64), counts, axis=0)
    f = ct[idx].view(np.float32).astype(np.float64)
    pa = b[a, 0:3] + _rotate(b[a, 4:8], f[:, 0:3])
    pb = b[bb, 0:3] + _rotate(b[bb, 4:8], f[:, 4:7])
    return -(normal * (pa - pb)).sum(1)


def world_statistics(bodies, manifolds, contacts, up=(0, 1, 0), gravity=-9.81, floor=-1.0):
    """Statistics over the dynamic bodies still above `floor` (a body whose lowest point is below it slid off the
    ground's edge and is counted as `fallen`, not in the heights and energies)."""
    b = np.asarray(bodies, np.float64)
    up = np.asarray(up, np.float64)
    low_all = lowest_points(b, up)
    dyn0 = b[:, 19] > 0
    dyn = dyn0 & (low_all >= floor)
    centre = b[dyn, 0:3] @ up
    low = low_all[dyn]
    kinetic, potential = energy(np.where(dyn[:, None] | ~dyn0[:, None], b, 0.0), up, gravity)
    depth = contact_depths(b, manifolds, contacts)
    speed = np.linalg.norm(b[dyn, 32:35], axis=1)
    out = {
        "bodies": int(dyn.sum()), "fallen": int((dyn0 & ~dyn).sum()),
        "centre_height": {"mean": float(centre.mean()), "min": float(centre.min()), "p50": float(np.median(centre)),
                          "max": float(centre.max())},
        "lowest_point": {"min": float(low.min()), "p1": float(np.percentile(low, 1)), "mean": float(low.mean())},
        "kinetic_energy": kinetic, "potential_energy": potential,
        "speed": {"mean": float(speed.mean()), "p99": float(np.percentile(speed, 99)), "max": float(speed.max())},
        "contacts": int(len(depth)),
    }
    if len(depth):
        out["penetration"] = {"max": float(depth.max()), "mean": float(depth.mean()),
                              "p99": float(np.percentile(depth, 99))}
    return out

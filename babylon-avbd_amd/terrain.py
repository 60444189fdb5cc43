"""Static terrain as convex pieces (src/native/terrain.js): every source triangle becomes a solid triangular prism, a
static hull body, so the ground keeps its concavities, holes and overhangs while the narrowphase only ever meets convex
shapes. Cooking is host-only and independent of any device.

* cookHeightfield (terrain.js:62-134): row-major heights over a (columns x rows) grid, two triangles per cell (a cell
  in `holes` gives none), each extruded down to `baseHeight` (default: two below the lowest height) along `upAxis`;
* cookStaticMesh (terrain.js:137-176): each non-degenerate triangle of a mesh extruded `thickness` against its winding
  normal (degenerate seam triangles are skipped);
* a piece is cooked in its own frame (centred on its corners, for precision far from the origin) from the fixed
  prism topology, wound outward whichever way the source triangle faces, and placed at its principal frame;
* createTerrainBodies (terrain.js:179-218): zero-mass hull bodies with the given friction, restitution and collision
  masks, never sleeping.
"""
import math

import numpy as np

from .hull import hullFromTriangles

# The closed prism over top corners 0, 1, 2 and their base copies 3, 4, 5 (terrain.js:13-15), and the same with every
# triangle's winding turned, for a source triangle that faces the other way
PRISM_FACES = [0, 1, 2, 3, 5, 4, 0, 3, 4, 0, 4, 1, 1, 4, 5, 1, 5, 2, 2, 5, 3, 2, 3, 0]
REVERSED_PRISM_FACES = [v for i in range(0, 24, 3) for v in (PRISM_FACES[i], PRISM_FACES[i + 2], PRISM_FACES[i + 1])]


def _finite(values, length, name):
    a = [float(x) for x in (values if values is not None else [])]
    if len(a) != length or not all(math.isfinite(x) for x in a):
        raise ValueError(f"{name} requires {length} finite values")
    return a


def _positive(value, name):
    if not (isinstance(value, (int, float)) and math.isfinite(value) and value > 0):
        raise ValueError(f"{name} must be positive and finite")
    return float(value)


def _piece(points):
    """One prism (top a, b, c then the base copies) as a hull at its place."""
    p = np.asarray(points, np.float64)
    center = p.mean(axis=0)
    a, b, c, back = p[0], p[1], p[2], p[3]
    outward = float(np.dot(np.cross(b - a, c - a), a - back))
    shape = hullFromTriangles((p - center).reshape(-1), PRISM_FACES if outward > 0 else REVERSED_PRISM_FACES)
    if shape is None:
        raise ValueError("Terrain contains a degenerate collision piece")
    return dict(shape=shape, position=tuple(float(v + center[k]) for k, v in enumerate(shape.center)),
                rotation=shape.rotation)


def cookHeightfield(heights=None, columns=None, rows=None, spacing=1, origin=(0, 0, 0), upAxis=1, baseHeight=None,
                    holes=None):
    """Static solid ground; heights are row-major; each cell keeps both triangles (terrain.js:62-134)."""
    for name, value in (("columns", columns), ("rows", rows)):
        if not (isinstance(value, int) and not isinstance(value, bool) and value >= 2):
            raise ValueError(f"{name} must be an integer >= 2")
    values = _finite(heights, columns * rows, "heights")
    offset = _finite(origin, 3, "origin")
    if upAxis not in (0, 1, 2):
        raise ValueError("upAxis must be 0, 1 or 2")
    step = [spacing, spacing] if isinstance(spacing, (int, float)) else _finite(spacing, 2, "spacing")
    for v in step:
        _positive(v, "spacing")
    base = min(values) - 2 if baseHeight is None else baseHeight
    if not (isinstance(base, (int, float)) and math.isfinite(base)) or any(h <= base for h in values):
        raise ValueError("baseHeight must be below every terrain height")
    cells = (columns - 1) * (rows - 1)
    if holes is not None and len(holes) != cells:
        raise ValueError("holes requires one value per terrain cell")
    axes = [k for k in range(3) if k != upAxis]
    pieces = []

    def point(x, z):
        q = list(offset)
        q[axes[0]] += x * step[0]
        q[axes[1]] += z * step[1]
        q[upAxis] += values[z * columns + x]
        return q

    for z in range(rows - 1):
        for x in range(columns - 1):
            if holes is not None and holes[z * (columns - 1) + x]:
                continue
            a, b, c, d = point(x, z), point(x + 1, z), point(x, z + 1), point(x + 1, z + 1)
            for top in ([a, b, c], [b, d, c]):
                bottom = []
                for q in top:
                    q = list(q)
                    q[upAxis] = offset[upAxis] + base
                    bottom.append(q)
                pieces.append(_piece(top + bottom))
    return dict(kind="heightfield", pieces=pieces, triangleCount=len(pieces))


def cookStaticMesh(positions=None, indices=None, thickness=0.1):
    """Triangle surfaces keep their concavities, holes and overhangs; each is backed along its winding normal
    (terrain.js:137-176)."""
    if positions is None or len(positions) % 3:
        raise ValueError("positions requires xyz triples")
    p = _finite(positions, len(positions), "positions")
    count = len(p) // 3
    if indices is None and count % 3:
        raise ValueError("Nonindexed positions require three vertices per triangle")
    ids = list(range(count)) if indices is None else list(indices)
    if not ids or len(ids) % 3 or not all(isinstance(i, (int, np.integer)) and 0 <= i < count for i in ids):
        raise ValueError("indices requires valid triangle vertex indices")
    _positive(thickness, "thickness")
    pieces = []
    for i in range(0, len(ids), 3):
        top = [np.array(p[3 * k:3 * k + 3]) for k in ids[i:i + 3]]
        a, b, c = top
        n = np.cross(b - a, c - a)
        length = float(np.linalg.norm(n))
        if length == 0:
            continue  # imported meshes often carry empty seam triangles
        pieces.append(_piece([*top, *[q - n * thickness / length for q in top]]))
    if not pieces:
        raise ValueError("Static mesh requires at least one nondegenerate triangle")
    return dict(kind="static-mesh", pieces=pieces, triangleCount=len(pieces))


def createTerrainBodies(scene, cooked, options=None):
    """Static hull bodies for cooked pieces (terrain.js:179-218); returns their handles."""
    options = options or {}
    if options.get("mass") not in (None, 0):
        raise ValueError("Terrain is static and requires zero mass")
    friction = options.get("friction", 0.6)
    restitution = options.get("restitution", 0)
    if not (isinstance(friction, (int, float)) and math.isfinite(friction) and friction >= 0):
        raise ValueError("friction must be nonnegative and finite")
    if not (isinstance(restitution, (int, float)) and math.isfinite(restitution) and 0 <= restitution <= 1):
        raise ValueError("restitution must be between 0 and 1")
    masks = {}
    for name in ("group", "collidesWith"):
        value = options.get(name, 0xFFFFFFFF)
        if not (isinstance(value, (int, np.integer)) and 0 <= value <= 0xFFFFFFFF):
            raise ValueError(f"{name} must be an unsigned 32-bit integer")
        masks[name] = int(value)
    if not cooked or not cooked.get("pieces"):
        raise ValueError("Terrain requires collision pieces")
    bodies = []
    for piece in cooked["pieces"]:
        bodies.append(scene.addHullShape(piece["shape"], density=0, friction=float(friction),
                                         position=piece["position"], rotation=piece["rotation"], allowSleep=False,
                                         restitution=float(restitution), group=masks["group"],
                                         collidesWith=masks["collidesWith"]))
    return bodies

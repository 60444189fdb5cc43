"""Convex hull shapes (reference/three-avbd/src/avbd3d/hull.ts: convexHull, hullFromTriangles, MAX_HULL_VERTICES; the
body builder shapes.ts:44-70). Cooking runs on the host, in the library (avbd_hull_from_points /
avbd_hull_from_triangles): no device is touched.

A HullShape holds the cooked hull in its principal frame (vertices, faces with outward unit normals and plane
offsets, edges with the two faces each separates), its mass properties per unit density, and the packed record the
narrowphase reads (header offsets relative to the record's first word; `record(at)` rebases them to where a scene's
hull buffer places it).
"""
import ctypes

import numpy as np

from . import _native

MAX_HULL_VERTICES = 32


class HullShape:
    """A cooked convex hull (hull.ts HullShape)."""

    def __init__(self, words, info):
        self._words = np.ascontiguousarray(words, np.uint32).reshape(-1, 4)
        self.volume = float(info.volume)
        self.moments = tuple(float(x) for x in info.moments)
        self.size = tuple(float(x) for x in info.size)
        self.radius = float(info.radius)
        self.center = tuple(float(x) for x in info.center)
        self.rotation = tuple(float(x) for x in info.rotation)
        h0, h1 = self._words[0], self._words[1]
        vs, nv, fs, nf = (int(x) for x in h0)
        es, ne, idx = int(h1[0]), int(h1[1]), int(h1[2])
        self.vertices = self._words[vs:vs + nv, 0:3].view(np.float32).copy()
        faces = []
        flat = self._words[idx:].reshape(-1)
        for k in range(nf):
            plane = self._words[fs + 2 * k].view(np.float32)
            first, count = int(self._words[fs + 2 * k + 1][0]), int(self._words[fs + 2 * k + 1][1])
            faces.append(dict(normal=tuple(float(x) for x in plane[0:3]), d=float(plane[3]),
                              verts=[int(v) for v in flat[first:first + count]]))
        self.faces = faces
        self.edges = self._words[es:es + ne].astype(np.int64).copy()

    @property
    def words(self):
        """Size of the packed record in 16-byte words."""
        return len(self._words)

    def record(self, at=0):
        """The packed record with its header at word `at` of the buffer it goes in ((words, 4) uint32)."""
        r = self._words.copy()
        r[0, 0] += at
        r[0, 2] += at
        r[1, 0] += at
        r[1, 2] += at
        return r


def _cook(fn, *args):
    lib = _native.load()
    info = _native.HullInfo()
    words = np.zeros((_native.HULL_MAX_WORDS, 4), np.uint32)
    code = getattr(lib, fn)(*args, 0, _native.ptr(words, ctypes.c_uint32), _native.HULL_MAX_WORDS, ctypes.byref(info))
    if code != 0:
        return None, lib.avbd_last_error().decode()
    return HullShape(words[:info.words], info), ""


def _points(points, name):
    p = np.ascontiguousarray(np.asarray(points, np.float64).reshape(-1))
    if p.size % 3 or not np.all(np.isfinite(p)):
        raise ValueError(f"{name} must contain finite xyz triples")
    return p


def convexHull(points):
    """The convex hull of xyz points (flat list or (n, 3)), at most MAX_HULL_VERTICES vertices (a bigger one is the hull
    of that many of its vertices, spread out), or None when the points span no volume."""
    p = _points(points, "Hull points")
    shape, _ = _cook("avbd_hull_from_points", _native.ptr(p, ctypes.c_double), p.size // 3)
    return shape


def hullFromTriangles(positions, indices):
    """A hull from a closed, outward-wound triangulation of a convex point set (coplanar neighbours merged into
    faces), or None when the triangles do not close up or enclose no volume."""
    p = _points(positions, "positions")
    idx = np.ascontiguousarray(np.asarray(indices, np.int64).reshape(-1))
    if idx.size % 3 or (idx.size and (idx.min() < 0 or idx.max() >= p.size // 3)):
        return None
    idx = idx.astype(np.uint32)
    shape, _ = _cook("avbd_hull_from_triangles", _native.ptr(p, ctypes.c_double), p.size // 3,
                     _native.ptr(idx, ctypes.c_uint32), idx.size)
    return shape

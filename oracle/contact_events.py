"""ORACLE — test infrastructure only (see math.hpp).

CPU restatement of the reference's contact events (src/vendor/three-avbd-lib/contacts.ts ContactWatch: touching,
ended; src/native/solver3D.js:252-311 watchContacts / readContactEvents / resetContactHistory) over the oracle world's
manifolds, points and bodies after each step, in FP32 in the order the HIP passes evaluate them. The sets are Python
sets: the reference's hash sets only matter when they overflow, which the tests do not reach.
"""
import numpy as np

BEGIN, END = 1, 2
f32 = np.float32


def _turn(q, v):
    """v + w·t + q×t with t = 2 q×v, component by component in f32."""
    u = q[:3]
    t = np.cross(u, v).astype(f32) * f32(2)
    return ((v + t * q[3]).astype(f32) + np.cross(u, t).astype(f32)).astype(f32)


class ContactEvents:
    def __init__(self, n_bodies):
        self.watch = np.ones(n_bodies, np.uint32)
        self.previous_watch = self.watch.copy()
        self.last = set()
        self.log = []

    def set_watched(self, indices, flags):
        for i, f in zip(indices, flags):
            self.watch[i] = (self.watch[i] & ~np.uint32(1)) | np.uint32(f)

    def reset(self, indices):
        for i in indices:
            self.watch[i] += np.uint32(2)

    def after_step(self, manifolds, contact_words, bodies, step, dt):
        """manifolds (m, 8) u32, contact words (c, 16) u32, bodies (n, 40) f32 of the step just taken."""
        cf = contact_words.view(np.float32)
        w, pw = self.watch, self.previous_watch
        now = set()
        for mf in manifolds:
            a, b, first, count = int(mf[0]), int(mf[1]), int(mf[2]), int(mf[3]) & 15
            if count == 0 or ((w[a] | w[b]) & 1) == 0:
                continue
            now.add((a, b))
            if (a, b) in self.last and w[a] == pw[a] and w[b] == pw[b]:
                continue
            pos, rot = bodies[a, 0:3].astype(f32), bodies[a, 4:8].astype(f32)
            point = np.zeros(3, f32)
            force = f32(0)
            for i in range(count):
                k = cf[first + i]
                point = (point + (_turn(rot, k[0:3]) + pos).astype(f32)).astype(f32)
                force = f32(force + k[12])
            point = (point / f32(count)).astype(f32)
            normal = mf[4:7].view(np.float32).copy()
            self.log.append((BEGIN, a, b, step, tuple(point.tolist()), float(f32(abs(force)) * f32(dt)),
                             tuple(normal.tolist())))
        for (a, b) in self.last:
            if w[a] != pw[a] or w[b] != pw[b]:
                continue
            if (a, b) not in now:
                self.log.append((END, a, b, step, (0.0, 0.0, 0.0), 0.0, (0.0, 0.0, 0.0)))
        self.last = now
        self.previous_watch = w.copy()

    def read(self):
        out, self.log = self.log, []
        return out

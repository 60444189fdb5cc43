"""Test-side loader of the CPU oracle (oracle/liboracle.so): the FP64 upstream restatement and the GPU-semantics
restatement (FP32 mirror / FP64). Only the tests, smoke() and bench.py's CPU baseline use it, as the checker."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None
u32p = ctypes.POINTER(ctypes.c_uint32)
i32p = ctypes.POINTER(ctypes.c_int32)
f32p = ctypes.POINTER(ctypes.c_float)
f64p = ctypes.POINTER(ctypes.c_double)
vp = ctypes.c_void_p


def _stale():
    if not os.path.exists(ORACLE_LIB):
        return True
    t = os.path.getmtime(ORACLE_LIB)
    srcs = [f for f in os.listdir(ORACLE_DIR) if f.endswith((".cpp", ".hpp")) or f == "Makefile"]
    return any(os.path.getmtime(os.path.join(ORACLE_DIR, f)) > t for f in srcs)


def build_oracle():
    if _stale():
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return ORACLE_LIB


def lib():
    global _lib
    if _lib is None:
        build_oracle()
        L = ctypes.CDLL(ORACLE_LIB)
        sig = {
            "ref_create": (vp, [ctypes.c_int, ctypes.c_int]),
            "ref_destroy": (None, [vp]),
            "ref_step": (None, [vp, ctypes.c_int]),
            "ref_body_count": (ctypes.c_int, [vp]),
            "ref_force_count": (ctypes.c_int, [vp]),
            "ref_read_bodies": (None, [vp, f64p]),
            "ref_joint_stats": (ctypes.c_int, [vp, f64p]),
            "ref_pick": (ctypes.c_int, [vp, f64p, f64p, f64p]),
            "ref_add_world_joint": (None, [vp, ctypes.c_int, f64p, f64p, ctypes.c_double, ctypes.c_double]),
            "ref_world_joint_error": (ctypes.c_double, [vp, ctypes.c_int, f64p, f64p]),
            "ref_box_mass": (None, [f64p, ctypes.c_double, f64p]),
            "gs_create": (vp, [ctypes.c_int, ctypes.c_uint32, f32p, u32p, ctypes.c_uint32, f32p, ctypes.c_uint32, f32p,
                               i32p, ctypes.c_uint32, u32p, u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]),
            "gs_destroy": (None, [vp]),
            "gs_set_brute_force": (None, [vp, ctypes.c_int]),
            "gs_set_coloring": (None, [vp, ctypes.c_uint32, ctypes.c_uint32]),
            "gs_set_pair_capacity": (None, [vp, ctypes.c_uint32]),
            "gs_set_hulls": (None, [vp, u32p, ctypes.c_uint32]),
            "gs_probe_hull_pair": (ctypes.c_int, [ctypes.c_int, f32p, f32p, u32p, ctypes.c_uint32, ctypes.c_int, u32p,
                                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                                  ctypes.POINTER(ctypes.c_double)]),
            "gs_overflow": (ctypes.c_uint32, [vp]),
            "gs_partial_step": (None, [vp, ctypes.c_int]),
            "cpu_create": (vp, [ctypes.c_uint32, f32p, ctypes.c_uint32, f32p, i32p, f32p, ctypes.c_int, ctypes.c_int]),
            "cpu_destroy": (None, [vp]),
            "cpu_step": (None, [vp, ctypes.c_int]),
            "cpu_set_threads": (None, [vp, ctypes.c_int]),
            "cpu_read_bodies": (None, [vp, f64p]),
            "cpu_counts": (ctypes.c_uint32, [vp, u32p]),
            "cpu_read_pairs": (ctypes.c_uint32, [vp, u32p]),
            "cpu_broadphase": (None, [vp]),
            "gs_capture_frames": (None, [vp, u32p, ctypes.c_uint32]),
            "gs_edit_joints": (None, [vp, u32p, ctypes.c_uint32]),
            "gs_set_state": (None, [vp, f32p, u32p, ctypes.c_uint32, u32p, ctypes.c_uint32, u32p, u32p]),
            "gs_step": (None, [vp, ctypes.c_int]),
            "gs_counts": (None, [vp, u32p]),
            "gs_read_bodies": (None, [vp, f64p]),
            "gs_read_joints": (None, [vp, f64p]),
            "gs_read_pairs": (None, [vp, u32p]),
            "gs_read_manifolds": (None, [vp, u32p]),
            "gs_read_contacts": (None, [vp, u32p, f64p]),
            "gs_read_contact_words": (None, [vp, u32p]),
            "gs_read_colors": (None, [vp, u32p]),
            "gs_read_adjacency": (ctypes.c_uint32, [vp, u32p, u32p]),
            "gs_read_refs": (None, [vp, f32p]),
            "gs_probe_box_contacts": (ctypes.c_int, [ctypes.c_int, f64p, f64p, f64p, f64p, ctypes.c_int, ctypes.c_int,
                                                     u32p, f64p, f64p]),
            "gs_probe_solve6": (None, [f64p, f64p, ctypes.c_double, f64p]),
            "gs_probe_pair_contacts": (ctypes.c_int, [ctypes.c_int, f32p, f32p, ctypes.c_int, u32p, f64p, f64p]),
            "gs_probe_priority": (ctypes.c_uint32, [ctypes.c_uint32]),
            "gs_apply_commands": (None, [vp, u32p, ctypes.c_uint32]),
            "gs_set_params": (None, [vp, ctypes.c_uint32, f32p]),
            "gs_set_iterations": (None, [vp, ctypes.c_int]),
            "gs_set_filters": (None, [vp, u32p, u32p, u32p, ctypes.c_uint32]),
            "gs_set_threads": (None, [ctypes.c_int]),
            "gs_set_restitution": (None, [vp, u32p, f32p, ctypes.c_uint32]),
            "gs_get_threads": (ctypes.c_int, []),
            "gs_enable_sleeping": (ctypes.c_int, [vp, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                                  ctypes.POINTER(ctypes.c_uint8)]),
            "gs_disable_sleeping": (None, [vp]),
            "gs_wake_all": (None, [vp]),
            "gs_wake_bodies": (None, [vp, u32p, ctypes.c_uint32]),
            "gs_set_sleep_enabled": (None, [vp, u32p, ctypes.POINTER(ctypes.c_uint8), ctypes.c_uint32]),
            "gs_sleep_stats": (None, [vp, ctypes.c_uint32, u32p]),
            "gs_read_sleep": (None, [vp, u32p]),
            "gs_set_sleep": (None, [vp, u32p, u32p]),
            "gs_cast": (None, [ctypes.c_uint32, f32p, u32p, f32p, ctypes.c_uint32, u32p, ctypes.c_uint32,
                               ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, f32p, u32p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def cast(bodies, filters, queries, ignore=(), mask=0xFFFFFFFF, membership=None, hulls=None):
    """The oracle's shape queries: body records (n, 40), filters (n, 2) u32 (group, collides-with), queries (m, 8)
    (origin, max distance, unit direction, radius), the hull buffer hull bodies name -> hits (m, 8): normal, distance
    or -1, index bits, 0, 0, 0."""
    b = np.ascontiguousarray(bodies, np.float32).reshape(-1, 40)
    f = np.ascontiguousarray(filters, np.uint32).reshape(-1, 2)
    q = np.ascontiguousarray(queries, np.float32).reshape(-1, 8)
    ig = np.ascontiguousarray(list(ignore) or [0], np.uint32)
    hits = np.zeros((len(q), 8), np.float32)
    hb = None if hulls is None or not len(hulls) else np.ascontiguousarray(hulls, np.uint32)
    lib().gs_cast(len(b), P(b, ctypes.c_float), P(f, ctypes.c_uint32), P(q, ctypes.c_float), len(q),
                  P(ig, ctypes.c_uint32), len(ignore), int(mask), int(membership or 0), int(membership is not None),
                  P(hits, ctypes.c_float), P(hb, ctypes.c_uint32))
    return hits


def set_threads(n):
    """Threads of the oracle's colour-parallel variant (1 = every loop in order; the results are the same)."""
    lib().gs_set_threads(int(n))


def get_threads():
    return int(lib().gs_get_threads())


def P(a, t):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


def no_collide(info, ignore):
    """(hi, lo, slot) triples as the engine builds them: joints with two bodies, then ignored pairs (slot ~0)."""
    out = []
    for j, (t, a, b, _) in enumerate(np.asarray(info).reshape(-1, 4)):
        if a >= 0:
            out.append((max(a, b), min(a, b), j))
    for a, b in np.asarray(ignore).reshape(-1, 2):
        out.append((max(a, b), min(a, b), 0xFFFFFFFF))
    return np.array(out, np.uint32).reshape(-1, 3)


def rebase_hulls(bodies, hulls, base):
    """A scene's body records and hull buffer moved to word `base` of a batch's buffer, as the engine concatenates
    them: hull shape codes and the four offsets of each named record's header grow by `base`."""
    b = np.array(bodies, np.float32, copy=True)
    hb = np.array(hulls, np.uint32, copy=True).reshape(-1, 4)
    codes = b[:, 39].astype(np.float64)
    named = sorted({int(c - 3) for c in codes if c >= 3})
    if base:
        b[codes >= 3, 39] = (codes[codes >= 3] + base).astype(np.float32)
        for h in named:
            hb[h, 0] += base
            hb[h, 2] += base
            hb[h + 1, 0] += base
            hb[h + 1, 2] += base
    return b, hb


def probe_hull_pair(recA, recB, hulls, face_bias=True, precision=32):
    """Contacts of one pair (any shapes, hulls from `hulls`): (features, points (k, 6), normal B->A, queries (8,):
    A-face sep, face, B-face sep, face, edge sep, valid, A edge, B edge; -3.4e38 where not run)."""
    a = np.ascontiguousarray(recA, np.float32)
    b = np.ascontiguousarray(recB, np.float32)
    h = np.ascontiguousarray(hulls, np.uint32).reshape(-1, 4)
    feats = np.zeros(8, np.uint32)
    pts = np.zeros((8, 6))
    normal = np.zeros(3)
    q = np.zeros(8)
    n = lib().gs_probe_hull_pair(precision, P(a, ctypes.c_float), P(b, ctypes.c_float), P(h, ctypes.c_uint32), len(h),
                                 int(face_bias), P(feats, ctypes.c_uint32), P(pts, ctypes.c_double),
                                 P(normal, ctypes.c_double), P(q, ctypes.c_double))
    return feats[:n], pts[:n], normal, q


class OracleWorld:
    """The GPU-semantics oracle over packed records (several scenes = a batch)."""

    def __init__(self, scenes, precision=32, color_cap=16, color_rounds=16, brute_force=False):
        L = lib()
        bodies, joints, info, ignore, filters, params, world = [], [], [], [], [], [], []
        base = 0
        hulls, hull_base = [], 0
        for k, s in enumerate(scenes):
            b, j, i, g, f = s.packed()
            b, hb = rebase_hulls(b, s.hull_buffer(), hull_base)
            hulls.append(hb)
            hull_base += len(hb)
            bodies.append(b)
            joints.append(j)
            i = i.copy()
            i[:, 1] = np.where(i[:, 1] >= 0, i[:, 1] + base, -1)
            i[:, 2] += base
            info.append(i)
            ignore.append(g.astype(np.int64) + base)
            filters.append(f if f is not None else np.full((len(b), 2), 0xFFFFFFFF, np.uint32))
            params.append(s.world_params())
            world.append(np.full(len(b), k, np.uint32))
            base += len(b)
        self.bodies0 = np.ascontiguousarray(np.concatenate(bodies), np.float32)
        self.joints0 = np.ascontiguousarray(np.concatenate(joints).reshape(-1, 32), np.float32)
        self.info = np.ascontiguousarray(np.concatenate(info).reshape(-1, 4), np.int32)
        ign = np.concatenate(ignore).reshape(-1, 2)
        self.nc = np.ascontiguousarray(no_collide(self.info, ign))
        self.filters = np.ascontiguousarray(np.concatenate(filters), np.uint32)
        self.params = np.ascontiguousarray(np.stack(params), np.float32)
        self.world = np.ascontiguousarray(np.concatenate(world), np.uint32)
        self.n = len(self.bodies0)
        self.nj = len(self.joints0)
        self.nw = len(scenes)
        self.precision = precision
        self.h = L.gs_create(precision, self.n, P(self.bodies0, ctypes.c_float), P(self.world, ctypes.c_uint32), self.nw,
                             P(self.params, ctypes.c_float), self.nj, P(self.joints0, ctypes.c_float),
                             P(self.info, ctypes.c_int32), len(self.nc), P(self.nc, ctypes.c_uint32),
                             P(self.filters, ctypes.c_uint32), color_cap, color_rounds,
                             int(scenes[0].params["iterations"]))
        self.hulls = np.ascontiguousarray(np.concatenate(hulls) if hull_base else np.zeros((0, 4), np.uint32))
        self._set_hulls()
        self._cap, self._rounds, self._iters = color_cap, color_rounds, int(scenes[0].params["iterations"])
        self._brute = brute_force
        if brute_force:
            L.gs_set_brute_force(self.h, 1)
        self._pair_capacity = 0
        L.gs_capture_frames(self.h, None, 0)  # as a new GPU world does from its starting poses
        self._restitution = None
        self._sleeping = None
        bouncy, base = [], 0
        for sc in scenes:  # bodies built with a restitution coefficient, as the HIP world applies them
            bouncy += [(base + i, e) for i, e in sorted(getattr(sc, "_restitution", {}).items())]
            base += sc.body_count
        if bouncy:
            self.set_restitution([i for i, _ in bouncy], [e for _, e in bouncy])

    def __del__(self):
        if getattr(self, "h", None):
            lib().gs_destroy(self.h)
            self.h = None

    def _set_hulls(self):
        if len(self.hulls):
            lib().gs_set_hulls(self.h, P(self.hulls, ctypes.c_uint32), len(self.hulls))

    def seed(self, state):
        """Start from a state read back from the HIP world (bodies, joints, refs, colours, contacts, steps)."""
        L = lib()
        # Rebuild with the given bodies and joints, then set the persistent state
        bodies = np.ascontiguousarray(state["bodies"], np.float32)
        joints = np.ascontiguousarray(state["joints"], np.float32).reshape(-1, 32)
        cc = ctypes.c_uint32
        old = self.h
        counts = np.zeros(8, np.uint32)
        L.gs_counts(old, P(counts, cc))
        self.h = L.gs_create(self.precision, self.n, P(bodies, ctypes.c_float), P(self.world, cc), self.nw,
                             P(self.params, ctypes.c_float), self.nj, P(joints, ctypes.c_float),
                             P(self.info, ctypes.c_int32), len(self.nc), P(self.nc, cc), P(self.filters, cc),
                             self._cap, self._rounds, self._iters)
        L.gs_destroy(old)
        self._set_hulls()
        if self._brute:
            L.gs_set_brute_force(self.h, 1)
        if self._pair_capacity:
            L.gs_set_pair_capacity(self.h, self._pair_capacity)
        refs = np.ascontiguousarray(state["refs"], np.float32)
        colors = np.ascontiguousarray(state["colors"], np.uint32)
        mf = np.ascontiguousarray(state["manifolds"], np.uint32)
        ct = np.ascontiguousarray(state["contacts"], np.uint32)
        steps = np.ascontiguousarray(state["steps"], np.uint32)
        L.gs_set_state(self.h, P(refs, ctypes.c_float), P(colors, cc), len(mf), P(mf, cc), len(ct), P(ct, cc),
                       P(steps, cc))
        if getattr(self, "_restitution", None) is not None:
            self.set_restitution(*self._restitution)
        if getattr(self, "_sleeping", None) is not None:
            self.enable_sleeping(**self._sleeping)
            if "sleep" in state:
                self.set_sleep_state(state["sleep"], state["sleep_globals"])

    # --- sleeping ---------------------------------------------------------------------------------------------------
    def enable_sleeping(self, speed=0.03, delay=0.5, islands=False, start_asleep=False, eligible=None):
        e = None if eligible is None else np.ascontiguousarray(eligible, np.uint8)
        rc = lib().gs_enable_sleeping(self.h, float(speed), float(delay), int(bool(islands)), int(bool(start_asleep)),
                                      P(e, ctypes.c_uint8))
        if rc != 0:
            raise ValueError("the oracle sleeps FP32 worlds only")
        self._sleeping = dict(speed=speed, delay=delay, islands=islands, start_asleep=False, eligible=eligible)

    def disable_sleeping(self):
        lib().gs_disable_sleeping(self.h)
        self._sleeping = None

    def wake_all(self):
        lib().gs_wake_all(self.h)

    def wake_bodies(self, indices):
        idx = np.ascontiguousarray(indices, np.uint32)
        lib().gs_wake_bodies(self.h, P(idx, ctypes.c_uint32), len(idx))

    def set_sleep_enabled(self, indices, flags):
        idx = np.ascontiguousarray(indices, np.uint32)
        f = np.ascontiguousarray(flags, np.uint8)
        lib().gs_set_sleep_enabled(self.h, P(idx, ctypes.c_uint32), P(f, ctypes.c_uint8), len(idx))

    def sleep_stats(self, k=0):
        out = np.zeros(2, np.uint32)
        lib().gs_sleep_stats(self.h, int(k), P(out, ctypes.c_uint32))
        return dict(sleeping=int(out[0]), wakeRequested=bool(out[1]))

    def sleep_state(self):
        """(n, 8) u32: mass bits, quiet bits, asleep, support, eligible, parent, wake, flags."""
        out = np.zeros((self.n, 8), np.uint32)
        lib().gs_read_sleep(self.h, P(out, ctypes.c_uint32))
        return out

    def set_sleep_state(self, states, globals_):
        st = np.ascontiguousarray(states, np.uint32).reshape(-1, 8)
        g = np.ascontiguousarray(globals_, np.uint32)
        lib().gs_set_sleep(self.h, P(st, ctypes.c_uint32), P(g, ctypes.c_uint32))

    def step(self, n=1):
        lib().gs_step(self.h, n)

    # --- constraint edits, as the GPU world applies them (the same record words) ----------------------------------
    def _edit(self, triples):
        t = np.ascontiguousarray(np.asarray(triples, np.uint32).reshape(-1, 3))
        if len(t):
            lib().gs_edit_joints(self.h, P(t, ctypes.c_uint32), len(t))

    def release_joints(self, slots):
        t = []
        for s in sorted(set(int(x) for x in slots)):
            if s < self.nj and self.info[s, 0] != 0:
                t += [(s, w, 0) for w in range(32)] + [(s, 32, 0), (s, 33, 0xFFFFFFFF), (s, 34, 0xFFFFFFFF)]
                self.info[s] = (0, -1, -1, 0)  # (the current types: a later seed() keeps them)
        self._edit(t)

    def disable_constraint(self, slot):
        self._edit([(slot, w, 0) for w in (0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 13, 14)])

    def set_world_anchor(self, slot, point):
        p = np.asarray(point, np.float32).view(np.uint32)
        self._edit([(slot, 24 + k, int(p[k])) for k in range(3)])

    def set_motor(self, slot, speed=None, max_torque=None):
        t = []
        if speed is not None:
            t.append((slot, 27, int(np.float32(speed).view(np.uint32))))
        if max_torque is not None:
            t.append((slot, 31, int(np.float32(max_torque).view(np.uint32))))
        self._edit(t)

    def set_spring_material(self, slot, break_strain, break_force):
        v = [0.0 if np.isinf(x) else x for x in (break_strain, break_force)]
        self._edit([(slot, 11, int(np.float32(v[0]).view(np.uint32))), (slot, 15, int(np.float32(v[1]).view(np.uint32)))])

    def set_angular_constraint(self, slot, kind, axis_a, axis_b, lo, hi):
        rec = np.zeros(32, np.float32)
        rec[0] = rec[1] = 1000.0
        rec[3] = 3e38
        for k, axis in ((24, axis_a), (28, axis_b)):
            a = np.asarray(axis, np.float32).astype(np.float64)
            rec[k:k + 3] = a / np.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])
        rec[27], rec[31] = lo, hi
        self._edit([(slot, w, int(rec.view(np.uint32)[w])) for w in range(32)] + [(slot, 32, int(kind))])
        self.info[slot, 0] = int(kind)
        s = np.array([slot], np.uint32)
        lib().gs_capture_frames(self.h, P(s, ctypes.c_uint32), 1)

    def partial_step(self, stage):
        """Stage 1: collision and topology only (contacts as warm-started); stage 2: also the warm starts."""
        lib().gs_partial_step(self.h, int(stage))

    def set_coloring(self, color_cap, color_rounds):
        """Colour cap and rounds from the next step on (also after a later seed())."""
        self._cap, self._rounds = int(color_cap), int(color_rounds)
        lib().gs_set_coloring(self.h, self._cap, self._rounds)

    def set_pair_capacity(self, capacity):
        """Pair storage like the GPU world's (0: unbounded): a step with more pairs overflows and is held."""
        self._pair_capacity = int(capacity)
        lib().gs_set_pair_capacity(self.h, self._pair_capacity)

    def counts(self):
        c = np.zeros(8, np.uint32)
        L = lib()
        L.gs_counts(self.h, P(c, ctypes.c_uint32))
        return dict(pairs=int(c[0]), manifolds=int(c[1]), contacts=int(c[2]), clashes=int(c[3]), colors=int(c[4]),
                    pivotFallbacks=int(c[7]), overflow=int(L.gs_overflow(self.h)))

    def bodies(self):
        out = np.zeros((self.n, 40))
        lib().gs_read_bodies(self.h, P(out, ctypes.c_double))
        return out

    def joints(self):
        out = np.zeros((self.nj, 32))
        if self.nj:
            lib().gs_read_joints(self.h, P(out, ctypes.c_double))
        return out

    def pairs(self):
        out = np.zeros((self.counts()["pairs"], 2), np.uint32)
        lib().gs_read_pairs(self.h, P(out, ctypes.c_uint32))
        return out

    def manifolds(self):
        out = np.zeros((self.counts()["manifolds"], 8), np.uint32)
        lib().gs_read_manifolds(self.h, P(out, ctypes.c_uint32))
        return out

    def contacts(self):
        n = self.counts()["contacts"]
        keys = np.zeros(n, np.uint32)
        vals = np.zeros((n, 16))
        lib().gs_read_contacts(self.h, P(keys, ctypes.c_uint32), P(vals, ctypes.c_double))
        return keys, vals

    def contact_words(self):
        out = np.zeros((self.counts()["contacts"], 16), np.uint32)
        lib().gs_read_contact_words(self.h, P(out, ctypes.c_uint32))
        return out

    def colors(self):
        out = np.zeros(self.n, np.uint32)
        lib().gs_read_colors(self.h, P(out, ctypes.c_uint32))
        return out

    def adjacency(self):
        L = lib()
        start = np.zeros(self.n + 1, np.uint32)
        m = L.gs_read_adjacency(self.h, P(start, ctypes.c_uint32), None)
        lst = np.zeros(m, np.uint32)
        L.gs_read_adjacency(self.h, P(start, ctypes.c_uint32), P(lst, ctypes.c_uint32))
        return start, lst

    def refs(self):
        out = np.zeros((self.n, 8), np.float32)
        lib().gs_read_refs(self.h, P(out, ctypes.c_float))
        return out

    def apply_commands(self, words):
        """Body commands (n x 12 u32 records), in submission order."""
        w = np.ascontiguousarray(words, np.uint32).reshape(-1, 12)
        lib().gs_apply_commands(self.h, P(w, ctypes.c_uint32), len(w))

    def set_params(self, k, dt, gravity, beta_lin, beta_ang, gamma, alpha, up, flags):
        """World k's parameters (its step counter is kept)."""
        v = np.array([dt, gravity, beta_lin, beta_ang, gamma, alpha, *up, 0], np.float32)
        v[9:10].view(np.uint32)[0] = flags
        self.params[k, :10] = v
        lib().gs_set_params(self.h, k, P(v, ctypes.c_float))

    def set_iterations(self, n):
        self._iters = int(n)
        lib().gs_set_iterations(self.h, int(n))

    def set_restitution(self, indices, values):
        idx = np.ascontiguousarray(indices, np.uint32)
        val = np.ascontiguousarray(values, np.float32)
        lib().gs_set_restitution(self.h, P(idx, ctypes.c_uint32), P(val, ctypes.c_float), len(idx))
        self._restitution = (idx, val)

    def set_filters(self, indices, groups, masks):
        idx = np.ascontiguousarray(indices, np.uint32)
        g = np.ascontiguousarray(groups, np.uint32)
        m = np.ascontiguousarray(masks, np.uint32)
        self.filters[idx, 0] = g
        self.filters[idx, 1] = m
        lib().gs_set_filters(self.h, P(idx, ctypes.c_uint32), P(g, ctypes.c_uint32), P(m, ctypes.c_uint32), len(idx))


class CpuPathWorld:
    """The FP64 restatement of the reference's CPU path (CpuSolver3D.step over the package's builder records: grid
    broadphase, face-biased box SAT, mass-penalty floor, sequential Gauss–Seidel newest first); threads > 1 runs its
    colour-parallel variant. Boxes, joints and springs only."""

    def __init__(self, scene, threads=1, bodies=None):
        L = lib()
        b, j, i, _, _ = scene.packed()
        if bodies is not None:
            b = np.ascontiguousarray(bodies, np.float32)
        self.bodies0 = np.ascontiguousarray(b, np.float32)
        self.joints0 = np.ascontiguousarray(j, np.float32).reshape(-1, 32)
        self.info = np.ascontiguousarray(i, np.int32).reshape(-1, 4)
        self.params = np.ascontiguousarray(scene.world_params(), np.float32)
        self.n = len(self.bodies0)
        self.h = L.cpu_create(self.n, P(self.bodies0, ctypes.c_float), len(self.joints0), P(self.joints0, ctypes.c_float),
                              P(self.info, ctypes.c_int32), P(self.params, ctypes.c_float),
                              int(scene.params["iterations"]), int(threads))
        if not self.h:
            raise ValueError("the CPU path restatement takes boxes, joints and springs only")

    def __del__(self):
        if getattr(self, "h", None):
            lib().cpu_destroy(self.h)
            self.h = None

    def step(self, n=1):
        lib().cpu_step(self.h, int(n))

    def set_threads(self, n):
        lib().cpu_set_threads(self.h, int(n))

    def bodies(self):
        out = np.zeros((self.n, 40))
        lib().cpu_read_bodies(self.h, P(out, ctypes.c_double))
        return out

    def counts(self):
        c = np.zeros(3, np.uint32)
        points = lib().cpu_counts(self.h, P(c, ctypes.c_uint32))
        return dict(manifolds=int(c[0]), joints=int(c[1]), springs=int(c[2]), contacts=int(points))

    def broadphase(self):
        """The broadphase alone: manifolds for the new pairs (not yet initialised)."""
        lib().cpu_broadphase(self.h)

    def pairs(self):
        L = lib()
        n = L.cpu_read_pairs(self.h, None)
        out = np.zeros((n, 2), np.uint32)
        L.cpu_read_pairs(self.h, P(out, ctypes.c_uint32))
        return out

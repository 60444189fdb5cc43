"""Worlds and pipelines: the reference's createSolverPipeline (src/native/pipeline.js:14-93) with a "hip" backend whose
worlds have the shape of its WebGPU worlds (src/native/solver3D.js:104-368, src/native/scenes.d.ts:317-454):
step / advance / reads / adapt / destroy. A batch world steps many independent scenes in one launch per phase.
"""
import ctypes
import math

import numpy as np

from . import _native, edits as _edits
from ._native import check, ptr
from .scene import AvbdScene3D


def _scene_desc(scene, keep):
    """A SceneDesc for a built AvbdScene3D; `keep` holds the arrays alive for the call."""
    bodies, joints, info, ignore, filters = scene.packed()
    keep.extend([bodies, joints, info, ignore, filters])
    d = _native.SceneDesc()
    d.body_count = len(bodies)
    d.bodies = ptr(bodies, ctypes.c_float)
    d.joint_count = len(joints)
    d.joints = ptr(joints, ctypes.c_float) if len(joints) else None
    d.joint_info = ptr(info, ctypes.c_int32) if len(joints) else None
    d.ignore_count = len(ignore)
    d.ignore_pairs = ptr(ignore, ctypes.c_uint32) if len(ignore) else None
    d.filters = ptr(filters, ctypes.c_uint32) if filters is not None else None
    p = scene.params
    d.dt = p["dt"]
    d.gravity = p["gravity"]
    d.up[:] = scene.up
    d.alpha = p["alpha"]
    d.gamma = p["gamma"]
    d.beta_lin = p["betaLin"]
    d.beta_ang = p["betaAng"]
    d.iterations = p["iterations"]
    d.flags = p["flags"]
    hulls = scene.hull_buffer()
    if len(hulls):
        keep.append(hulls)
        d.hull_words = len(hulls)
        d.hulls = ptr(hulls, ctypes.c_uint32)
    return d


def _options(options):
    o = _native.Options()
    o.color_cap = int(options.get("colorCap", 0))
    o.color_rounds = int(options.get("colorRounds", 0))
    cap = options.get("capacity", {}) or {}
    o.pair_capacity = int(cap.get("pairs", 0))
    o.use_graph = 1 if options.get("graph", False) else 0
    o.device = int(options.get("device", 0))
    o.adaptive = 1 if options.get("adaptive", False) else 0
    bp = options.get("broadphase", "auto")
    if bp not in ("auto", "grid", "hploc"):
        raise ValueError(f"Unknown broadphase: {bp}")
    o.broadphase = {"auto": 0, "grid": 1, "hploc": 2}[bp]
    o.bvh_rebuild_interval = int((options.get("bvh") or {}).get("rebuildInterval", 0))
    solver = options.get("solver", "auto")
    if solver not in ("auto", "colors", "dataflow", "worlds"):
        raise ValueError(f"Unknown solver: {solver}")
    o.solver = {"auto": 0, "colors": 1, "dataflow": 2, "worlds": 3}[solver]
    block = int(options.get("worldBlock", 0))
    if block not in (0, 128, 256, 512, 1024):
        raise ValueError(f"worldBlock must be 128, 256, 512 or 1024, not {block}")
    o.world_block = block
    share = int(options.get("dataflowShare", 0))
    if not 0 <= share <= 1024:
        raise ValueError(f"dataflowShare must be between 0 and 1024, not {share}")
    o.dataflow_share = share
    rounds = options.get("coloringRounds", "auto")
    if rounds not in ("auto", "launches", "listed"):
        raise ValueError(f"Unknown coloringRounds: {rounds}")
    o.coloring_rounds = {"auto": 0, "launches": 1, "listed": 2}[rounds]
    return o


class HipWorld:
    """A world (or a batch of worlds) on the GPU. Methods mirror AvbdGpuSolver3D (scenes.d.ts:317-454)."""

    backend = "hip"

    def __init__(self, scenes, options=None):
        options = options or {}
        lib = _native.load()
        scenes = list(scenes)
        if not scenes:
            raise ValueError("An AVBD scene builder is required")
        for s in scenes:
            if not isinstance(s, AvbdScene3D):
                raise TypeError("An AVBD scene builder is required")
            s.assertEditable()
        keep = []
        descs = (_native.SceneDesc * len(scenes))(*[_scene_desc(s, keep) for s in scenes])
        handle = ctypes.c_void_p()
        check(lib.avbd_batch_create(descs, len(scenes), ctypes.byref(_options(options)), ctypes.byref(handle)))
        self._lib = lib
        self._h = handle
        self.destroyed = False
        self.scenes = scenes
        for s in scenes:
            s.created = True
        self.world_first_body = np.cumsum([0] + [s.body_count for s in scenes])[:-1]
        self.world_first_joint = np.cumsum([0] + [len(s._joints) for s in scenes])[:-1]
        p = scenes[0].params
        self.params = dict(p, up=list(scenes[0].up))
        self.timeAccumulator = 0.0
        n = ctypes.c_uint32()
        check(lib.avbd_world_body_count(self._h, ctypes.byref(n)))
        self.bodyCount = n.value
        check(lib.avbd_world_joint_count(self._h, ctypes.byref(n)))
        self.jointCount = n.value
        self.worldCount = len(scenes)
        self._dts = [float(sc.params["dt"]) for sc in scenes]
        self._pending = _edits.CommandStaging()
        # Bodies built with a restitution coefficient (scenes.js:684-687 applies them after creation the same way)
        bouncy = [(int(first) + i, e) for first, sc in zip(self.world_first_body, scenes)
                  for i, e in sorted(sc._restitution.items())]
        if bouncy:
            self.setRestitution([i for i, _ in bouncy], [e for _, e in bouncy])
        self._masses = np.concatenate([np.concatenate(s._chunks)[:, 19] if s._chunks else np.zeros(0, np.float32)
                                       for s in scenes]).astype(np.float32)
        # Sleeping (scenes.js:689-693): asked for in the options, or a body was built with allowSleep true
        no_sleep = [int(first) + i for first, sc in zip(self.world_first_body, scenes) for i in sorted(sc._no_sleep)]
        if no_sleep:
            flags = np.zeros(len(no_sleep), np.uint8)
            idx = np.array(no_sleep, np.uint32)
            check(lib.avbd_world_set_sleep_enabled(self._h, ptr(idx, ctypes.c_uint32), ptr(flags, ctypes.c_uint8),
                                                   len(idx)))
        sleeping = options.get("sleeping")
        if sleeping or any(s.sleepOptIn for s in scenes):
            self.enableSleeping(sleeping if isinstance(sleeping, dict) else {})

    # --- lifetime ---------------------------------------------------------------------------------------------------

    def _alive(self):
        if self.destroyed:
            raise RuntimeError("The solver has been destroyed")

    def _ready(self):
        """Alive, with queued edits submitted (the reference flushes them before steps and reads)."""
        self._alive()
        self.flushEdits()

    def destroy(self):
        if self.destroyed:
            return
        check(self._lib.avbd_world_destroy(self._h))
        self.destroyed = True

    def __del__(self):
        try:
            if not self.destroyed and _native._lib is not None:
                self._lib.avbd_world_destroy(self._h)
                self.destroyed = True
        except Exception:
            pass

    # --- stepping ---------------------------------------------------------------------------------------------------

    def step(self, n=1):
        """Submit n steps without waiting (solver3D.js:350-359); queued edits go first."""
        self._ready()
        check(self._lib.avbd_world_step(self._h, int(n)))

    def advance(self, elapsed, maxSubSteps=6, maxFrameTime=0.05):
        """Fixed-step accumulator; returns the steps submitted (runtime.js:255-272)."""
        self._alive()
        if not math.isfinite(elapsed) or elapsed < 0:
            raise ValueError("elapsed must be finite nonnegative seconds")
        if int(maxSubSteps) != maxSubSteps or maxSubSteps < 1 or not math.isfinite(maxFrameTime) or maxFrameTime <= 0:
            raise ValueError("maxSubSteps must be a positive integer and maxFrameTime positive")
        dt = self.params["dt"]
        self.timeAccumulator = min(maxSubSteps * dt, self.timeAccumulator + min(maxFrameTime, elapsed))
        count = 0
        while self.timeAccumulator + 1e-12 >= dt and count < maxSubSteps:
            self.step()
            self.timeAccumulator = max(0.0, self.timeAccumulator - dt)
            count += 1
        return count

    def stepTimed(self, n=1):
        """Run n steps and wait; returns each step's device milliseconds (HIP events around each step)."""
        self._ready()
        ms = np.zeros(int(n), np.float32)
        check(self._lib.avbd_world_step_timed(self._h, int(n), ptr(ms, ctypes.c_float)))
        return ms

    def stepTimedBegin(self, n=1):
        """Enqueue n event-bracketed steps without waiting (pair with stepTimedEnd)."""
        self._ready()
        check(self._lib.avbd_world_step_timed_begin(self._h, int(n)))
        self._timed = int(n)

    def stepTimedEnd(self):
        """Wait for the steps of the last stepTimedBegin; returns their device milliseconds."""
        self._alive()
        n = getattr(self, "_timed", 0)
        ms = np.zeros(max(n, 1), np.float32)
        got = ctypes.c_uint32()
        check(self._lib.avbd_world_step_timed_end(self._h, ptr(ms, ctypes.c_float), len(ms), ctypes.byref(got)))
        self._timed = 0
        return ms[:got.value]

    def synchronize(self):
        self._ready()
        check(self._lib.avbd_world_synchronize(self._h))

    # --- edits between steps (solver3D.js:130-213, runtime.js:45-175) --------------------------------------------

    def liveIndex(self, index):
        self._alive()
        if isinstance(index, bool) or not isinstance(index, (int, np.integer)) or not 0 <= index < self.bodyCount:
            raise ValueError("A live 3D GPU body index is required")
        return int(index)

    def _world_of(self, index):
        return int(np.searchsorted(self.world_first_body, index, side="right") - 1)

    def _dt_of(self, index):
        return self._dts[self._world_of(index)]

    def queueBodyCommand(self, index, kind, value, point=None, hasPoint=False):
        self._pending.append(self.liveIndex(index), kind, value, point, hasPoint)
        return self

    def setLinearVelocity(self, index, v):
        return self.queueBodyCommand(index, _edits.LINEAR_VELOCITY, _edits.components(v, 3, "velocity"))

    def setAngularVelocity(self, index, v):
        return self.queueBodyCommand(index, _edits.ANGULAR_VELOCITY, _edits.components(v, 3, "angular velocity"))

    def applyImpulse(self, index, v, point=None):
        p = None if point is None else _edits.components(point, 3, "point")
        return self.queueBodyCommand(index, _edits.IMPULSE, _edits.components(v, 3, "impulse"), p, p is not None)

    def applyForce(self, index, v, point=None):
        """A force over the next step: the impulse v × dt (dt of the body's world)."""
        p = None if point is None else _edits.components(point, 3, "point")
        index = self.liveIndex(index)
        return self.queueBodyCommand(index, _edits.FORCE, _edits.components(v, 3, "force") + [self._dt_of(index)], p,
                                     p is not None)

    def applyAngularImpulse(self, index, v):
        return self.queueBodyCommand(index, _edits.ANGULAR_IMPULSE, _edits.components(v, 3, "angular impulse"))

    def applyTorque(self, index, v):
        index = self.liveIndex(index)
        return self.queueBodyCommand(index, _edits.TORQUE, _edits.components(v, 3, "torque") + [self._dt_of(index)])

    def teleport(self, index, position, rotation=None):
        p = _edits.components(position, 3, "position")
        q = None if rotation is None else _edits.unit_rotation(rotation)
        return self.queueBodyCommand(index, _edits.TELEPORT, p, q, q is not None)

    def setPoses(self, indices, positions, rotations=None):
        """Packed teleports (solver3D.js:188-206): unique ascending u32 slots, f32 xyz positions, optional f32 xyzw."""
        self._alive()
        staged = _edits.CommandStaging()
        _edits.stage_poses(indices, positions, rotations, self.bodyCount, staged)
        self._pending.extend(staged)
        return self

    def editBodies(self, edits):
        """editBodies (runtime.js:45-175): the whole batch is checked before anything is queued; per record a teleport
        precedes velocity replacements, impulses and forces; group/collidesWith edits become setFilters."""
        self._alive()
        staged = _edits.CommandStaging()
        properties = _edits.stage_edits(edits, self.bodyIndex, self._dt_of, staged)
        self._pending.extend(staged)
        for index, group, collides, restitution, allow in properties:
            if group is not None:
                self.setFilters([index], [group], [collides])
            if restitution is not None:
                self.setRestitution([index], [restitution])
            if allow is not None:
                self.setSleepEnabled(index, allow)
        return self

    # --- sleeping (worldSleeping.js; solver3D.js:317-343; appGpuSolver3D.js:869-889) -------------------------------

    def enableSleeping(self, options=None):
        """enableSleeping({speedThreshold, timeThreshold, islands, startAsleep}): bodies quiet and supported for
        timeThreshold seconds sleep as fixed supports until an impact, an edit or a request wakes them. islands (default:
        worlds of 4096 bodies or more) wakes and sleeps connected groups together instead of the whole world;
        startAsleep (island mode) takes one step that puts still, grounded groups to sleep at once."""
        self._alive()
        o = dict(options or {})
        unknown = set(o) - {"speedThreshold", "timeThreshold", "islands", "startAsleep"}
        if unknown:
            raise ValueError(f"Unknown sleep option: {sorted(unknown)[0]}")
        speed, delay = o.get("speedThreshold", 0.03), o.get("timeThreshold", 0.5)
        ok = all(isinstance(v, (int, float)) and not isinstance(v, bool) and math.isfinite(v) and v > 0
                 for v in (speed, delay))
        if not ok:
            raise ValueError("Sleep speedThreshold and timeThreshold must be positive and finite")
        if o.get("startAsleep") is not None and not isinstance(o["startAsleep"], bool):
            raise ValueError("startAsleep must be a boolean")
        if o.get("islands") is not None and not isinstance(o["islands"], bool):
            raise ValueError("Sleep islands must be a boolean")
        if self.sleeping:
            return self
        islands = -1 if o.get("islands") is None else int(o["islands"])
        start = bool(o.get("startAsleep", False))
        self.flushEdits()
        check(self._lib.avbd_world_enable_sleeping(self._h, float(speed), float(delay), islands, int(start)))
        if start and self.sleepIslands:
            self.step()
        return self

    @property
    def sleeping(self):
        return bool(self._lib.avbd_world_sleeping(self._h, None)) if not self.destroyed else False

    @property
    def sleepIslands(self):
        mode = ctypes.c_int(0)
        on = self._lib.avbd_world_sleeping(self._h, ctypes.byref(mode))
        return bool(on and mode.value)

    def disableSleeping(self):
        """Every body wakes with its mass back (poses kept) and sleeping stops."""
        self._alive()
        self.flushEdits()
        check(self._lib.avbd_world_disable_sleeping(self._h))
        return self

    def wakeAll(self):
        self._alive()
        self.flushEdits()
        check(self._lib.avbd_world_wake_all(self._h))
        return self

    def wakeBodies(self, indices):
        """wakeBody for each index: held awake through the next step; in island mode their islands wake."""
        self._alive()
        idx = np.array([self.liveIndex(i) for i in np.atleast_1d(indices)], np.uint32)
        self.flushEdits()
        check(self._lib.avbd_world_wake_bodies(self._h, ptr(idx, ctypes.c_uint32), len(idx)))
        return self

    def setSleepEnabled(self, index, enabled):
        """Sleep eligibility of one body; allowing it enables sleeping; every body wakes."""
        self._alive()
        if not isinstance(enabled, bool):
            raise ValueError("Sleep flag must be boolean")
        idx = np.array([self.liveIndex(index)], np.uint32)
        flag = np.array([int(enabled)], np.uint8)
        self.flushEdits()
        check(self._lib.avbd_world_set_sleep_enabled(self._h, ptr(idx, ctypes.c_uint32), ptr(flag, ctypes.c_uint8), 1))
        return self

    def readSleepStats(self, world=0):
        """{sleeping, wakeRequested} of one world after the last step."""
        self._ready()
        out = np.zeros(2, np.uint32)
        check(self._lib.avbd_world_sleep_stats(self._h, int(world), ptr(out, ctypes.c_uint32)))
        return {"sleeping": int(out[0]), "wakeRequested": bool(out[1])}

    # --- shape queries (appGpuSolver3D.js:306-332; shapeQueries.js:99-258) ------------------------------------------

    def raycast(self, origin, direction, options=None):
        """The nearest hit of one ray, or None (raycastAll of one ray)."""
        o = dict(options or {})
        return self.raycastAll([{"origin": origin, "direction": direction, "maxDistance": o.get("maxDistance")}], o)[0]

    def raycastAll(self, rays, options=None):
        """Rays {origin, direction, maxDistance?} against every body's current pose: per ray None or
        {index, distance, normal, center, point}. Options: ignore (up to 16 bodies), collidesWith (a body is seen when
        its group & collidesWith != 0), membership (also its collides-with mask & membership != 0), includeTriggers,
        maxDistance (the batch default)."""
        return self._cast([dict(r, radius=0.0) for r in rays], options)

    def sphereCast(self, origin, radius, direction, options=None):
        o = dict(options or {})
        return self.sphereCastAll([{"origin": origin, "radius": radius, "direction": direction,
                                    "maxDistance": o.get("maxDistance")}], o)[0]

    def sphereCastAll(self, casts, options=None):
        """Spheres {origin, radius, direction, maxDistance?} swept along their directions; results as raycastAll's,
        `point` being the touching point on the body (center − radius · normal)."""
        return self._cast(casts, options)

    def _cast(self, queries, options):
        self._ready()
        o = dict(options or {})
        ignore = list(o.get("ignore", []))
        mask = o.get("collidesWith", 0xFFFFFFFF)
        membership = o.get("membership")
        include = o.get("includeTriggers", True)
        if isinstance(mask, bool) or not isinstance(mask, (int, np.integer)) or not 0 <= mask <= 0xFFFFFFFF:
            raise ValueError("Query mask must be an unsigned 32-bit integer")
        if not isinstance(include, bool):
            raise ValueError("includeTriggers must be boolean")
        if membership is not None and (isinstance(membership, bool) or not isinstance(membership, (int, np.integer))
                                       or not 0 <= membership <= 0xFFFFFFFF):
            raise ValueError("Query membership must be an unsigned 32-bit integer")
        if len(ignore) > 16 or any(isinstance(i, bool) or not isinstance(i, (int, np.integer)) or not
                                   0 <= i < self.bodyCount for i in ignore):
            raise ValueError("Queries can ignore up to 16 live body indices")
        if len(queries) > 65535:
            raise ValueError("Too many queries in one batch")
        batch_max = o.get("maxDistance")
        rows, checked = [], []
        for q in queries:
            origin = [float(x) for x in q["origin"]]
            direction = [float(x) for x in q["direction"]]
            radius = q.get("radius", 0.0)
            radius = 0.0 if radius is None else float(radius)
            md = q.get("maxDistance")
            md = (batch_max if batch_max is not None else 3e38) if md is None else md
            md = float(md)
            if len(origin) != 3 or len(direction) != 3 or not all(math.isfinite(x) for x in origin + direction) \
                    or math.hypot(*direction) == 0:
                raise ValueError("A query requires 3 finite components and a nonzero direction")
            if not (math.isfinite(radius) and radius >= 0 and math.isfinite(md) and md >= 0):
                raise ValueError("Query radius and maximum distance must be finite and nonnegative")
            length = math.hypot(*direction)
            unit = [x / length for x in direction]
            rows.append(origin + [md] + unit + [radius])
            checked.append((origin, unit, radius))
        if not rows:
            return []
        data = np.ascontiguousarray(rows, np.float32)
        hits = np.zeros((len(rows), 8), np.float32)
        ign = np.ascontiguousarray(ignore, np.uint32)
        opts = _native.QueryOptions(ign.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) if len(ign) else None,
                                    len(ign), int(mask), int(membership or 0), int(membership is not None))
        check(self._lib.avbd_world_cast(self._h, ptr(data, ctypes.c_float), len(rows), ctypes.byref(opts),
                                        ptr(hits, ctypes.c_float)))
        out = []
        for (origin, unit, radius), h in zip(checked, hits):
            if h[3] < 0:
                out.append(None)
                continue
            d = float(h[3])
            normal = h[0:3].astype(float).tolist()
            center = [origin[k] + d * unit[k] for k in range(3)]
            out.append({"index": int(h[4:5].view(np.uint32)[0]), "distance": d, "normal": normal, "center": center,
                        "point": [center[k] - radius * normal[k] for k in range(3)]})
        return out

    # --- contact events (solver3D.js:252-311; contacts.ts ContactWatch) --------------------------------------------

    def watchContacts(self, maxPairs=8192, maxEvents=4096, indices=None):
        """Bounded contact-event storage, made on the first call (every body watched); indices: only these report."""
        self._alive()
        for n in (maxPairs, maxEvents):
            if isinstance(n, bool) or not isinstance(n, (int, np.integer)) or n < 1:
                raise ValueError("Contact capacities must be positive integers")
        if indices is not None and not isinstance(indices, (list, tuple)):
            raise ValueError("Watched indices must be an array")
        idx = [self.liveIndex(i) for i in (indices or [])]
        self.flushEdits()
        check(self._lib.avbd_world_watch_contacts(self._h, int(maxPairs), int(maxEvents)))
        self._event_capacity = getattr(self, "_event_capacity", int(maxEvents))
        if indices is not None:
            selected = set(idx)
            all_ = np.arange(self.bodyCount, dtype=np.uint32)
            flags = np.array([int(i in selected) for i in range(self.bodyCount)], np.uint8)
            check(self._lib.avbd_world_set_watched(self._h, ptr(all_, ctypes.c_uint32), ptr(flags, ctypes.c_uint8),
                                                   self.bodyCount))
        return self

    def readContactEvents(self):
        """{events: [{type 'begin'|'end', a, b, step, point, impulse, normal, isTrigger}], dropped, full}: the events
        since the last read, in step order, clearing them."""
        self._ready()
        if not hasattr(self, "_event_capacity"):
            self.watchContacts()
        cap = self._event_capacity
        out = np.zeros((cap, 12), np.uint32)
        count, written, full = (np.zeros(1, np.uint32) for _ in range(3))
        check(self._lib.avbd_world_read_contact_events(self._h, ptr(out, ctypes.c_uint32), cap,
                                                       ptr(count, ctypes.c_uint32), ptr(written, ctypes.c_uint32),
                                                       ptr(full, ctypes.c_uint32)))
        rows = out[:int(count[0])]
        f = rows.view(np.float32)
        events = [{"type": "begin" if r[0] == 1 else "end", "kind": int(r[0]), "a": int(r[1]), "b": int(r[2]),
                   "step": int(r[3]), "point": f[k, 4:7].tolist(), "impulse": float(f[k, 7]),
                   "normal": f[k, 8:11].tolist(), "isTrigger": False} for k, r in enumerate(rows)]
        events.sort(key=lambda e: e["step"])
        return {"events": events, "dropped": int(written[0] - count[0]), "full": bool(full[0])}

    def resetContactHistory(self, indices):
        """New lifetimes for sorted unique slots: their pairs begin again and do not end."""
        self._alive()
        idx = np.ascontiguousarray(indices, np.uint32)
        self.flushEdits()
        check(self._lib.avbd_world_reset_contact_history(self._h, ptr(idx, ctypes.c_uint32), len(idx)))
        return self

    def readSleep(self):
        """Every body's sleep state as an (n, 8) u32 array (mass and quiet as f32 bits, asleep, support, eligible,
        parent, wake, flags) and the (worlds, 4) sleep globals."""
        self._ready()
        states = np.zeros((self.bodyCount, 8), np.uint32)
        globals_ = np.zeros((self.worldCount, 4), np.uint32)
        check(self._lib.avbd_world_read_sleep(self._h, ptr(states, ctypes.c_uint32), ptr(globals_, ctypes.c_uint32)))
        return states, globals_

    def setRestitution(self, indices, values):
        """setRestitution (appGpuSolver3D.js:455-471) for one body (index, value) or many: coefficients in [0, 1],
        from the next step on; while any is above 0 each step ends with the bounce passes (src/gpu/restitution.js)."""
        self._alive()
        if np.ndim(indices) == 0:
            indices, values = [indices], [values]
        idx = np.array([self.liveIndex(i) for i in indices], np.uint32)
        val = np.array([_edits.coefficient(v) for v in values], np.float32)
        if len(idx) != len(val):
            raise ValueError("Each body index needs a restitution value")
        self.flushEdits()  # commands queued before keep their place ahead of the change
        check(self._lib.avbd_world_set_restitution(self._h, ptr(idx, ctypes.c_uint32), ptr(val, ctypes.c_float),
                                                   len(idx)))
        return self

    def getRestitution(self):
        self._alive()
        out = np.zeros(self.bodyCount, np.float32)
        check(self._lib.avbd_world_get_restitution(self._h, ptr(out, ctypes.c_float)))
        return out

    def setFilters(self, indices, groups, collidesWith):
        """setFilters (appGpuSolver3D.js:958-983): collision group and mask per body, from the next step on."""
        self._alive()
        if not (len(indices) == len(groups) == len(collidesWith)):
            raise ValueError("Each body index needs a group and collision mask")
        idx = np.ascontiguousarray([self.liveIndex(i) for i in indices], np.uint32)
        g = np.ascontiguousarray([_edits.mask(int(v) if isinstance(v, np.integer) else v) for v in groups], np.uint32)
        m = np.ascontiguousarray([_edits.mask(int(v) if isinstance(v, np.integer) else v) for v in collidesWith],
                                 np.uint32)
        check(self._lib.avbd_world_set_filters(self._h, ptr(idx, ctypes.c_uint32), ptr(g, ctypes.c_uint32),
                                               ptr(m, ctypes.c_uint32), len(idx)))
        return self

    def flushEdits(self):
        """Submit queued body commands without stepping (solver3D.js:208-214)."""
        self._alive()
        if self._pending.count:
            words = np.ascontiguousarray(self._pending.records())
            self._pending.clear()
            check(self._lib.avbd_world_enqueue_commands(self._h, ptr(words, ctypes.c_uint32), len(words)))
        return self

    def getParams(self, world=0):
        """Step parameters of world `world` of the batch (GpuParams3D, solver.ts:350)."""
        self._alive()
        p = _native.Params()
        check(self._lib.avbd_world_get_params(self._h, int(world), ctypes.byref(p)))
        return dict(dt=p.dt, gravity=p.gravity, up=list(p.up), alpha=p.alpha, gamma=p.gamma, betaLin=p.beta_lin,
                    betaAng=p.beta_ang, iterations=p.iterations, flags=p.flags)

    def setParams(self, world=0, **changes):
        """Change step parameters (dt, gravity, up, alpha, gamma, betaLin, betaAng, iterations, flags) of world
        `world` from the next step on; iterations apply to every world of a batch."""
        self._ready()
        unknown = set(changes) - {"dt", "gravity", "up", "alpha", "gamma", "betaLin", "betaAng", "iterations", "flags"}
        if unknown:
            raise ValueError(f"Unknown parameters: {sorted(unknown)}")
        cur = self.getParams(world)
        cur.update(changes)
        p = _native.Params()
        p.dt = cur["dt"]
        p.gravity = cur["gravity"]
        p.up[:] = [float(v) for v in cur["up"]]
        p.alpha = cur["alpha"]
        p.gamma = cur["gamma"]
        p.beta_lin = cur["betaLin"]
        p.beta_ang = cur["betaAng"]
        p.iterations = int(cur["iterations"])
        p.flags = int(cur["flags"])
        check(self._lib.avbd_world_set_params(self._h, int(world), ctypes.byref(p)))
        self._dts[int(world)] = float(np.float32(p.dt))
        if int(world) == 0:
            self.params.update(dt=self._dts[0], gravity=p.gravity, iterations=p.iterations)
        return self

    # --- reads ----------------------------------------------------------------------------------------------------

    def readBodies(self):
        """All bodies, 40 f32 each, in slot order (solver.ts:1421-1423)."""
        self._ready()
        out = np.zeros((self.bodyCount, 40), np.float32)
        check(self._lib.avbd_world_read_bodies(self._h, 0, self.bodyCount, ptr(out, ctypes.c_float)))
        return out

    def writeBodies(self, first, records):
        self._ready()
        records = np.ascontiguousarray(records, np.float32).reshape(-1, 40)
        check(self._lib.avbd_world_write_bodies(self._h, int(first), len(records), ptr(records, ctypes.c_float)))
        self._masses[int(first):int(first) + len(records)] = records[:, 19]  # the awake masses from now on

    def readSelectedBodies(self, indices, posesOnly=False):
        """readSelectedBodies (solver3D.js:223-230): 40 or (posesOnly) 8 f32 per body."""
        self._ready()
        idx = np.ascontiguousarray(indices, np.uint32)
        out = np.zeros((len(idx), 8 if posesOnly else 40), np.float32)
        check(self._lib.avbd_world_read_selected(self._h, ptr(idx, ctypes.c_uint32), len(idx), 1 if posesOnly else 0,
                                                 ptr(out, ctypes.c_float)))
        return out

    def readJoints(self):
        self._ready()
        out = np.zeros((self.jointCount, 32), np.float32)
        if self.jointCount:
            check(self._lib.avbd_world_read_joints(self._h, 0, self.jointCount, ptr(out, ctypes.c_float)))
        return out

    def writeJoints(self, first, records):
        self._ready()
        records = np.ascontiguousarray(records, np.float32).reshape(-1, 32)
        check(self._lib.avbd_world_write_joints(self._h, int(first), len(records), ptr(records, ctypes.c_float)))

    def readCounters(self):
        """{pairs, contacts, manifolds, joints, overflow, clashes, colors, sleeping} (world.js:672-692)."""
        self._alive()
        c = _native.Counters()
        check(self._lib.avbd_world_read_counters(self._h, ctypes.byref(c)))
        return dict(pairs=c.pairs, contacts=c.contacts, manifolds=c.manifolds, joints=c.joints, overflow=c.overflow,
                    clashes=c.clashes, colors=c.colors,
                    sleeping=self.readSleepStats()["sleeping"] if self.sleeping else 0, pivotFallbacks=c.pivot_fallbacks,
                    solveAborts=c.solve_aborts)

    def readPairs(self):
        """The last step's broadphase pairs (a > b) as an (n, 2) u32 array, in emission order."""
        self._alive()
        n = ctypes.c_uint32()
        check(self._lib.avbd_world_read_pairs(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros((n.value, 2), np.uint32)
        check(self._lib.avbd_world_read_pairs(self._h, ptr(out, ctypes.c_uint32), n.value, ctypes.byref(n)))
        return out

    def readContactRecords(self):
        """The last step's manifolds (m, 8) u32 and contacts (c, 16) u32, contacts packed densely per manifold."""
        self._ready()
        m = ctypes.c_uint32()
        c = ctypes.c_uint32()
        check(self._lib.avbd_world_read_contacts(self._h, None, 0, ctypes.byref(m), None, 0, ctypes.byref(c)))
        mf = np.zeros((m.value, 8), np.uint32)
        ct = np.zeros((c.value, 16), np.uint32)
        check(self._lib.avbd_world_read_contacts(self._h, ptr(mf, ctypes.c_uint32), m.value, ctypes.byref(m),
                                                 ptr(ct, ctypes.c_uint32), c.value, ctypes.byref(c)))
        return mf, ct

    def readContactList(self):
        """readContactList (solver.ts:1435-1469): one dict per contact point."""
        mf, ct = self.readContactRecords()
        out = []
        for row in mf:
            a, b, first, cw = (int(x) for x in row[:4])
            normal = row[4:7].view(np.float32).tolist()
            for c in ct[first:first + (cw & 15)]:
                key = int(c[3])
                f = c.view(np.float32)
                out.append(dict(a=a, b=b, feature=key & 0x7FFFFFFF, stick=bool(key & 0x80000000),
                                rA=f[0:3].tolist(), rB=f[4:7].tolist(), pen=f[8:11].tolist(), lam=f[12:15].tolist(),
                                normal=normal))
        return out

    def readColors(self):
        self._alive()
        out = np.zeros(self.bodyCount, np.uint32)
        check(self._lib.avbd_world_read_colors(self._h, ptr(out, ctypes.c_uint32)))
        return out

    def readAdjacency(self):
        self._alive()
        n = ctypes.c_uint32()
        start = np.zeros(self.bodyCount + 1, np.uint32)
        check(self._lib.avbd_world_read_adjacency(self._h, ptr(start, ctypes.c_uint32), None, 0, ctypes.byref(n)))
        lst = np.zeros(n.value, np.uint32)
        check(self._lib.avbd_world_read_adjacency(self._h, ptr(start, ctypes.c_uint32), ptr(lst, ctypes.c_uint32),
                                                  n.value, ctypes.byref(n)))
        return start, lst

    def readRefs(self):
        self._alive()
        out = np.zeros((self.bodyCount, 8), np.float32)
        check(self._lib.avbd_world_read_refs(self._h, ptr(out, ctypes.c_float)))
        return out

    def readSteps(self):
        self._alive()
        out = np.zeros(self.worldCount, np.uint32)
        check(self._lib.avbd_world_read_steps(self._h, ptr(out, ctypes.c_uint32)))
        return out

    def readState(self):
        """Everything the next step depends on: bodies, joints, refs, colours, last contacts, step counters."""
        mf, ct = self.readContactRecords()
        return dict(bodies=self.readBodies(), joints=self.readJoints(), refs=self.readRefs(), colors=self.readColors(),
                    manifolds=mf, contacts=ct, steps=self.readSteps())

    def writeState(self, state):
        """seedFrom (solver.ts:905-948): bodies, joints and the persistent state of readState()."""
        self._ready()
        self.writeBodies(0, state["bodies"])
        if self.jointCount:
            self.writeJoints(0, state["joints"])
        refs = np.ascontiguousarray(state["refs"], np.float32)
        colors = np.ascontiguousarray(state["colors"], np.uint32)
        mf = np.ascontiguousarray(state["manifolds"], np.uint32)
        ct = np.ascontiguousarray(state["contacts"], np.uint32)
        steps = np.ascontiguousarray(state["steps"], np.uint32)
        check(self._lib.avbd_world_write_state(self._h, ptr(refs, ctypes.c_float), ptr(colors, ctypes.c_uint32),
                                               len(mf), ptr(mf, ctypes.c_uint32) if len(mf) else None, len(ct),
                                               ptr(ct, ctypes.c_uint32) if len(ct) else None,
                                               ptr(steps, ctypes.c_uint32)))

    # --- policy ---------------------------------------------------------------------------------------------------

    def setColoring(self, colorCap, colorRounds=0):
        self._alive()
        check(self._lib.avbd_world_set_coloring(self._h, int(colorCap), int(colorRounds)))

    # --- constraint edits (appGpuSolver3D.js:500-563, 810-837, 922-957; solver.ts:1138-1204) -----------------------

    def releaseJoints(self, slots):
        """Release constraint slots: they stop acting and their bodies collide again."""
        self._alive()
        s = np.ascontiguousarray(list(slots), np.uint32)
        check(self._lib.avbd_world_release_joints(self._h, ptr(s, ctypes.c_uint32) if len(s) else None, len(s)))

    def disableConstraint(self, slot):
        self._alive()
        check(self._lib.avbd_world_disable_constraint(self._h, int(slot)))

    def setWorldAnchor(self, slot, point):
        self._alive()
        p = np.ascontiguousarray(point, np.float32)
        if p.shape != (3,):
            raise ValueError("A world joint and finite xyz anchor are required")
        check(self._lib.avbd_world_set_world_anchor(self._h, int(slot), ptr(p, ctypes.c_float)))

    def setMotor(self, slot, speed=None, maxTorque=None):
        self._alive()
        sp = None if speed is None else np.array([speed], np.float32)
        mt = None if maxTorque is None else np.array([maxTorque], np.float32)
        check(self._lib.avbd_world_set_motor(self._h, int(slot), ptr(sp, ctypes.c_float), ptr(mt, ctypes.c_float)))

    def setSpringMaterial(self, slot, breakStrain=math.inf, breakForce=math.inf):
        self._alive()
        check(self._lib.avbd_world_set_spring_material(self._h, int(slot), float(breakStrain), float(breakForce)))

    def setAngularConstraint(self, slot, type, axisA=(0, 1, 0), axisB=(0, 1, 0), speed=1.0, maxTorque=100.0,
                             minAngle=-math.pi, maxAngle=math.pi):
        """The slot becomes a motor ("motor": speed, maxTorque) or a limit ("limit": minAngle, maxAngle)."""
        self._alive()
        if type not in ("motor", "limit"):
            raise ValueError("Invalid angular constraint type")
        a = np.ascontiguousarray(axisA, np.float32)
        b = np.ascontiguousarray(axisB, np.float32)
        lo, hi = (speed, maxTorque) if type == "motor" else (minAngle, maxAngle)
        check(self._lib.avbd_world_set_angular_constraint(self._h, int(slot), 3 if type == "motor" else 4,
                                                          ptr(a, ctypes.c_float), ptr(b, ctypes.c_float), float(lo),
                                                          float(hi)))

    def jointInfo(self, slot):
        """(type, body A or -1, body B, 0) of a constraint slot."""
        self._alive()
        out = (ctypes.c_int32 * 4)()
        check(self._lib.avbd_world_joint_info(self._h, int(slot), out))
        return tuple(out)

    def getColoring(self):
        """{colorCap, colorRounds, pairCapacity} in use (as setColoring and adapt() leave them)."""
        self._alive()
        cap, rounds, pairs = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        check(self._lib.avbd_world_get_coloring(self._h, ctypes.byref(cap), ctypes.byref(rounds), ctypes.byref(pairs)))
        return dict(colorCap=cap.value, colorRounds=rounds.value, pairCapacity=pairs.value)

    def adapt(self, counters=None):
        """Grow storage and the colour cap from the last step's counters (solver.ts:1484-1519)."""
        self._alive()
        changed = ctypes.c_int()
        check(self._lib.avbd_world_adapt(self._h, ctypes.byref(changed)))
        return bool(changed.value)

    def profileStep(self):
        """One step with HIP events per phase and per primal/dual launch (ms)."""
        self._ready()
        out = np.zeros(13, np.float32)
        check(self._lib.avbd_world_profile_step(self._h, ptr(out, ctypes.c_float)))
        keys = ("total", "collision", "adjacency", "coloring", "warmStart", "primal", "dual", "velocities",
                "primalLaunches", "dualLaunches", "broadphase", "narrowphase", "pairSearch")
        return {k: float(v) for k, v in zip(keys, out)}

    def deviceBytes(self):
        v = ctypes.c_uint64()
        check(self._lib.avbd_world_device_bytes(self._h, ctypes.byref(v)))
        return v.value

    def solverKind(self):
        """Which solve the steps run: 1 colour launches, 2 dataflow, 3 one workgroup per world."""
        v = ctypes.c_uint32()
        check(self._lib.avbd_world_solver_kind(self._h, ctypes.byref(v)))
        return v.value

    def gpuIndex(self, i):
        """Slot of the scene builder's body i (bodies keep builder order)."""
        return self.liveIndex(i)

    def bodyIndex(self, body):
        """The slot of a handle of this world, or of a body index (runtime.js:326-345)."""
        self._alive()
        if isinstance(body, HipBody):
            if body.solver is not self:
                raise ValueError("Body belongs to another solver")
            return body.gpuIndex
        return self.liveIndex(body)

    def body(self, body):
        """The handle of a body (runtime.js:168-237), one per slot."""
        index = self.bodyIndex(body)
        handles = self.__dict__.setdefault("_handles", {})
        if index not in handles:
            handles[index] = HipBody(self, index)
        return handles[index]

    def readBodyState(self, index):
        """Position, rotation, linear and angular velocity of one body (queued edits applied first)."""
        index = self.liveIndex(index)
        r = self.readSelectedBodies([index])[0]
        mass = float(self._masses[index])
        return {"position": r[0:3].tolist(), "rotation": r[4:8].tolist(), "linearVelocity": r[32:35].tolist(),
                "angularVelocity": r[36:39].tolist(), "mass": mass, "effectiveMass": float(r[19]),
                "sleeping": mass > 0 and r[19] == 0}

    def setVelocities(self, indices, linear=None, angular=None):
        """setVelocities(indices, {linear, angular}) (runtime.js:356-378): xyz per body, the batch checked whole."""
        self._alive()
        if linear is None and angular is None:
            raise ValueError("Provide linear and/or angular velocities")
        idx = [self.liveIndex(i) for i in indices]

        def take(values, name):
            if values is None:
                return None
            a = np.asarray(values, np.float64).reshape(-1)
            if a.size != 3 * len(idx) or not np.isfinite(a).all():
                raise ValueError(f"{name} velocities must be finite xyz per body")
            return a.reshape(-1, 3)
        lin, ang = take(linear, "Linear"), take(angular, "Angular")
        staged = _edits.CommandStaging()
        for k, i in enumerate(idx):
            if lin is not None:
                staged.append(i, _edits.LINEAR_VELOCITY, lin[k].tolist())
            if ang is not None:
                staged.append(i, _edits.ANGULAR_VELOCITY, ang[k].tolist())
        self._pending.extend(staged)
        return self

    def setLinearVelocities(self, indices, values):
        return self.setVelocities(indices, linear=values)

    def setAngularVelocities(self, indices, values):
        return self.setVelocities(indices, angular=values)



class HipBody:
    """A body handle (runtime.js:168-237 NativeBody): chainable edits of one slot and readState(). Trigger bodies are
    not built on this backend: turning one on raises, turning it off is accepted."""

    def __init__(self, solver, index):
        self.solver = solver
        self.index = index

    @property
    def gpuIndex(self):
        return self.solver.liveIndex(self.index)

    def setLinearVelocity(self, v):
        self.solver.setLinearVelocity(self.gpuIndex, v)
        return self

    def setAngularVelocity(self, v):
        self.solver.setAngularVelocity(self.gpuIndex, v)
        return self

    def applyImpulse(self, v, point=None):
        self.solver.applyImpulse(self.gpuIndex, v, point)
        return self

    def applyForce(self, v, point=None):
        self.solver.applyForce(self.gpuIndex, v, point)
        return self

    def applyAngularImpulse(self, v):
        self.solver.applyAngularImpulse(self.gpuIndex, v)
        return self

    def applyTorque(self, v):
        self.solver.applyTorque(self.gpuIndex, v)
        return self

    def teleport(self, position, rotation=None):
        self.solver.teleport(self.gpuIndex, position, rotation)
        return self

    def setCollisionGroups(self, group, collidesWith=0xFFFFFFFF):
        self.solver.setFilters([self.gpuIndex], [group], [collidesWith])
        return self

    def setTrigger(self, enabled):
        self.gpuIndex
        if enabled:
            raise ValueError("Trigger bodies are not available on the hip backend")
        return self

    def setRestitution(self, value):
        self.solver.setRestitution(self.gpuIndex, value)
        return self

    def setSleepEnabled(self, enabled):
        self.solver.setSleepEnabled(self.gpuIndex, enabled)
        return self

    def wakeUp(self):
        self.solver.wakeBodies([self.gpuIndex])
        return self

    def readState(self):
        return self.solver.readBodyState(self.gpuIndex)


def createSolverPipeline(backend="hip", device=0):
    """createSolverPipeline({backend}) (pipeline.js:14-93) with the "hip" backend."""
    if backend != "hip":
        raise ValueError(f"Unsupported solver backend: {backend}")
    return SolverPipeline(device)


class SolverPipeline:
    def __init__(self, device=0):
        self.backend = "hip"
        self.device = int(device)
        self.worlds = set()
        self.destroyed = False

    def createSolver(self, scene, options=None):
        if self.destroyed:
            raise RuntimeError("The solver pipeline has been destroyed")
        opts = dict(options or {})
        opts.setdefault("device", self.device)
        w = HipWorld([scene], opts)
        self.worlds.add(w)
        return w

    def createBatch(self, scenes, options=None):
        """Independent worlds stepped together, one launch per phase."""
        if self.destroyed:
            raise RuntimeError("The solver pipeline has been destroyed")
        opts = dict(options or {})
        opts.setdefault("device", self.device)
        w = HipWorld(scenes, opts)
        self.worlds.add(w)
        return w

    def _ready(self):
        """Alive, with queued edits submitted (the reference flushes them before steps and reads)."""
        self._alive()
        self.flushEdits()

    def destroy(self):
        if self.destroyed:
            return
        for w in list(self.worlds):
            w.destroy()
        self.worlds.clear()
        self.destroyed = True

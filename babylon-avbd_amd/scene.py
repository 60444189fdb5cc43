"""Scene builder: the reference's AvbdScene3D (src/native/scenes.js:107-702) over packed records.

Bodies are returned as integer handles (their index, which is also their slot in the world: bodies are stored in
builder order). Records are packed as GpuSolver3D.pack / writeJoint do (reference/three-avbd/src/avbd3d/gpu/
solver.ts:794-850) in the layouts of include/avbd.h. Mass properties follow ref/body.ts:38-47 (boxes),
avbd3d/shapes.ts:28-36 (spheres) and src/gpu/capsuleShape.js:4-41 (capsules).
"""
import bisect
import math

import numpy as np

from . import _native

BIG = 3.0e38
T_JOINT = 1
T_SPRING = 2
T_MOTOR = 3
T_LIMIT = 4
SHAPE_BOX = 0.0
SHAPE_SPHERE = 1.0
SHAPE_CAPSULE = -1.0
SHAPE_HULL = 3.0  # + the word of the body's hull record in the scene's hull buffer


def _finite(x):
    """Infinity is never stored on the device: unbounded values are ±BIG (solver.ts `finite`)."""
    if x == math.inf:
        return BIG
    if x == -math.inf:
        return -BIG
    return float(x)


def _vector(v, n, name):
    a = [float(x) for x in v]
    if len(a) != n or not all(math.isfinite(x) for x in a):
        raise ValueError(f"{name} requires {n} finite components")
    return a


def _positive(v, name, zero=False):
    v = float(v)
    if not math.isfinite(v) or (v < 0 if zero else v <= 0):
        raise ValueError(f"{name} must be {'nonnegative' if zero else 'positive'} and finite")
    return v


def _rotate(q, v):
    """ref/math.ts rotate (x, y, z, w quaternion)."""
    ux, uy, uz, w = q
    tx = (uy * v[2] - uz * v[1]) * 2
    ty = (uz * v[0] - ux * v[2]) * 2
    tz = (ux * v[1] - uy * v[0]) * 2
    return [v[0] + tx * w + (uy * tz - uz * ty), v[1] + ty * w + (uz * tx - ux * tz), v[2] + tz * w + (ux * ty - uy * tx)]


def box_mass(size, density):
    """Box mass, principal moments and bounding radius (ref/body.ts:38-47)."""
    sx, sy, sz = size
    m = sx * sy * sz * density
    moment = [((sy * sy + sz * sz) / 12) * m, ((sx * sx + sz * sz) / 12) * m, ((sx * sx + sy * sy) / 12) * m]
    hx, hy, hz = sx * 0.5, sy * 0.5, sz * 0.5
    return m, moment, math.sqrt(hx * hx + hy * hy + hz * hz)


def capsule_mass(radius, height, density):
    """Solid capsule along local Y of full tip-to-tip height (src/gpu/capsuleShape.js:28-38)."""
    length = height - 2 * radius
    r2, l2 = radius * radius, length * length
    cylinder = math.pi * radius * radius * length * density
    caps = (4 / 3) * math.pi * (r2 * radius) * density
    axial = (cylinder * r2) / 2 + (caps * 2 * r2) / 5
    radial = cylinder * (l2 / 12 + r2 / 4) + caps * ((2 * r2) / 5 + l2 / 4 + (3 * length * radius) / 8)
    return cylinder + caps, [radial, axial, radial], height / 2


def body_record(position, rotation, size, mass, moment, radius, friction, shape, velocity=(0, 0, 0),
                angular_velocity=(0, 0, 0)):
    """One 40-float body record as GpuSolver3D.pack writes a freshly built body."""
    r = np.zeros(40, np.float32)
    r[0:3] = position
    r[3] = friction
    r[4:8] = rotation
    r[15] = 1.0  # initialRot (0, 0, 0, 1)
    r[16:19] = size
    r[19] = mass
    r[20:23] = moment
    r[23] = radius
    r[31] = 1.0  # inertialRot
    r[32:35] = velocity
    r[35] = 0.0  # previous velocity along up: a new body's is zero
    r[36:39] = angular_velocity
    r[39] = shape
    return r


_HEIGHTFIELD_KEYS = {"heights", "columns", "rows", "spacing", "origin", "upAxis", "baseHeight", "holes"}
_MESH_KEYS = {"positions", "indices", "thickness"}


def _unit(v):
    a = _vector(v, 3, "axis")
    n = math.sqrt(sum(x * x for x in a))
    if not (n >= 1e-8):
        raise ValueError("Angular axis requires three finite, nonzero components")
    return [x / n for x in a]


def _break_limits(breakStrain, breakForce):
    """springMaterial (src/gpu/springMaterial.js:4-15): positive or Infinity; stored 0 for Infinity."""
    out = []
    for name, v in (("breakStrain", breakStrain), ("breakForce", breakForce)):
        if v != math.inf and not (isinstance(v, (int, float)) and math.isfinite(v) and v > 0):
            raise ValueError(f"{name} must be positive or Infinity")
        out.append(0.0 if v == math.inf else float(v))
    return out


def _angular_record(kind, axisA, axisB, speed=1.0, maxTorque=100.0, minAngle=-math.pi, maxAngle=math.pi):
    """The record setAngularConstraint writes (appGpuSolver3D.js:500-517; angularConstraintOptions,
    angularConstraints.js:84-120): row penalties 1000, stiffness 3e38, unit axes, speed/torque bound or angles."""
    if not (isinstance(speed, (int, float)) and math.isfinite(speed) and isinstance(maxTorque, (int, float))
            and math.isfinite(maxTorque) and maxTorque >= 0):
        raise ValueError("Motor speed must be finite and maxTorque nonnegative")
    if not (math.isfinite(minAngle) and math.isfinite(maxAngle) and minAngle >= -math.pi and maxAngle <= math.pi
            and minAngle <= maxAngle):
        raise ValueError("Hinge angles must be ordered within -PI..PI radians")
    rec = np.zeros(32, np.float32)
    rec[0] = rec[1] = 1000.0
    rec[3] = 3e38
    rec[24:27] = _unit(axisA)
    rec[28:31] = _unit(axisB)
    rec[27] = speed if kind == "motor" else minAngle
    rec[31] = maxTorque if kind == "motor" else maxAngle
    return rec


class AvbdScene3D:
    """Portable scene builder (scenes.js:107-273). Defaults: Y up, gravity -10, dt 1/60, 10 iterations."""

    def __init__(self, gravity=-10.0, iterations=10, timeStep=1 / 60, up=(0, 1, 0), alpha=0.99, gamma=0.999,
                 betaLin=10000.0, betaAng=100.0, flags=_native.FLAGS_DEFAULT):
        if not math.isfinite(gravity):
            raise ValueError("gravity must be finite")
        if int(iterations) != iterations or iterations < 1:
            raise ValueError("iterations must be an integer")
        self.params = dict(dt=_positive(timeStep, "timeStep"), gravity=float(gravity), iterations=int(iterations),
                           alpha=float(alpha), gamma=float(gamma), betaLin=float(betaLin), betaAng=float(betaAng),
                           flags=int(flags))
        self.up = _vector(up, 3, "up")
        if abs(math.sqrt(sum(x * x for x in self.up)) - 1) > 1e-5:
            raise ValueError("up must be a unit vector")
        self._chunks = []         # (n, 40) float32 record blocks
        self._starts = []         # first body of each block
        self._count = 0
        self._filters = {}        # body -> (group, collidesWith)
        self._restitution = {}    # body -> restitution coefficient above 0
        self._no_sleep = set()    # bodies built with allowSleep false
        self.sleepOptIn = False   # a body was built with allowSleep true: its world sleeps (scenes.js:233-240, 689)
        self._joints = []         # (record[32], info[4])
        self._hulls = {}          # id(HullShape) -> (shape, header word): one record per shape, shared by its bodies
        self._hull_words = 0
        self._ignore = []
        self.created = False

    # --- bodies ---------------------------------------------------------------------------------------------------

    @property
    def body_count(self):
        return self._count

    def assertEditable(self):
        if self.created:
            raise RuntimeError("Build this scene before createSolver; use the solver's edit methods afterwards")

    def _add(self, rec, options):
        self.assertEditable()
        if options.get("rotation") is not None:
            q = _vector(options["rotation"], 4, "rotation")
            l = math.sqrt(sum(x * x for x in q))
            if not l:
                raise ValueError("rotation must be nonzero")
            rec[4:8] = [x / l for x in q]
        if options.get("mass") is not None:
            mass = _positive(options["mass"], "mass", True)
            if rec[19] == 0 and mass > 0:
                raise ValueError("A positive mass needs positive density")
            ratio = mass / rec[19] if rec[19] > 0 else 0.0
            rec[19] = mass
            rec[20:23] = rec[20:23] * np.float32(ratio)
        if options.get("velocity") is not None:
            rec[32:35] = _vector(options["velocity"], 3, "velocity")
        if options.get("angularVelocity") is not None:
            rec[36:39] = _vector(options["angularVelocity"], 3, "angularVelocity")
        restitution = options.get("restitution")
        if restitution is not None:
            r = float(restitution) if isinstance(restitution, (int, float)) else math.nan
            if not (math.isfinite(r) and 0 <= r <= 1):
                raise ValueError("restitution must be between 0 and 1")
        group = options.get("group", 0xFFFFFFFF)
        mask = options.get("collidesWith", 0xFFFFFFFF)
        for m in (group, mask):
            if int(m) != m or m < 0 or m > 0xFFFFFFFF:
                raise ValueError("Collision masks must be unsigned 32-bit integers")
        allow = options.get("allowSleep")
        if allow is not None and not isinstance(allow, bool):
            raise ValueError("allowSleep must be a boolean")
        index = self._count
        if allow is False:
            self._no_sleep.add(index)
        self.sleepOptIn = self.sleepOptIn or allow is True
        if group != 0xFFFFFFFF or mask != 0xFFFFFFFF:
            self._filters[index] = (int(group), int(mask))
        if restitution is not None and r > 0:
            self._restitution[index] = r
        self._starts.append(self._count)
        self._chunks.append(rec[None, :])
        self._count += 1
        return index

    def createSolver(self, pipeline, options=None):
        """scene.createSolver(pipeline, options) (scenes.js:652-702): a world of this scene on the pipeline's device."""
        if not hasattr(pipeline, "createSolver"):
            raise ValueError("A solver pipeline is required")
        return pipeline.createSolver(self, options)

    def addBox(self, size, **options):
        """A centred box with full XYZ extents; zero density makes it static (scenes.js:131-144)."""
        s = _vector(size, 3, "size")
        for v in s:
            _positive(v, "size")
        density = _positive(options.get("density", 1), "density", True)
        friction = _positive(options.get("friction", 0.6), "friction", True)
        pos = _vector(options.get("position", (0, 0, 0)), 3, "position")
        m, moment, radius = box_mass(s, density)
        return self._add(body_record(pos, (0, 0, 0, 1), s, m, moment, radius, friction, SHAPE_BOX), options)

    def addSphere(self, radius, **options):
        """A solid sphere (scenes.js:145-159; shapes.ts:28-36)."""
        r = _positive(radius, "radius")
        density = _positive(options.get("density", 1), "density", True)
        friction = _positive(options.get("friction", 0.6), "friction", True)
        pos = _vector(options.get("position", (0, 0, 0)), 3, "position")
        m = (4 / 3) * math.pi * r * r * r * density if density > 0 else 0.0
        inertia = 0.4 * m * r * r
        return self._add(body_record(pos, (0, 0, 0, 1), (2 * r, 2 * r, 2 * r), m, (inertia,) * 3, r, friction,
                                     SHAPE_SPHERE), options)

    def addCapsule(self, radius, height, **options):
        """An analytic capsule along local Y; height is the full tip-to-tip extent (scenes.js:181-197)."""
        r = float(radius)
        h = float(height)
        if not (math.isfinite(r) and r > 0 and math.isfinite(h) and h >= 2 * r):
            raise ValueError("Capsule height must be at least twice its positive radius")
        density = _positive(options.get("density", 1), "density", True)
        friction = _positive(options.get("friction", 0.6), "friction", True)
        pos = _vector(options.get("position", (0, 0, 0)), 3, "position")
        m, moment, bound = capsule_mass(r, h, density)
        return self._add(body_record(pos, (0, 0, 0, 1), (2 * r, h, 2 * r), m, moment, bound, friction, SHAPE_CAPSULE),
                         options)

    def addHull(self, points, **options):
        """A convex hull of xyz points, its principal frame at `position` (its centre of mass), as scenes.js:160-179:
        the points' hull (at most 32 vertices) of the given density; `rotation` turns the principal frame."""
        from .hull import convexHull
        self.assertEditable()
        flat = np.asarray(points, np.float64).reshape(-1)
        if flat.size % 3 or not np.all(np.isfinite(flat)):
            raise ValueError("Hull points must contain finite xyz triples")
        shape = convexHull(flat)
        if shape is None:
            raise ValueError("A convex hull requires finite points spanning a volume")
        return self.addHullShape(shape, **options)

    def addHullShape(self, shape, **options):
        """A body of a cooked HullShape (shapes.ts:50-70 `hull`): size, mass, moments and radius from the shape."""
        self.assertEditable()
        if shape.vertices.shape[0] > 32 or len(shape.faces) > 255 or len(shape.edges) > 255:
            raise ValueError(f"hull: {shape.vertices.shape[0]} vertices, {len(shape.faces)} faces, "
                             f"{len(shape.edges)} edges (at most 32 vertices, 255 faces and edges)")
        density = _positive(options.get("density", 1), "density", True)
        friction = _positive(options.get("friction", 0.6), "friction", True)
        pos = _vector(options.get("position", (0, 0, 0)), 3, "position")
        entry = self._hulls.get(id(shape))
        if entry is None:
            entry = (shape, self._hull_words)
            self._hulls[id(shape)] = entry
            self._hull_words += shape.words
        at = entry[1]
        m = shape.volume * density if density > 0 else 0.0
        moment = [x * max(density, 0.0) for x in shape.moments]
        rec = body_record(pos, (0, 0, 0, 1), shape.size, m, moment, shape.radius, friction, SHAPE_HULL + at)
        if rec[39] != np.float32(SHAPE_HULL + at):
            raise ValueError("hull buffer beyond 2^24 words")
        return self._add(rec, options)

    def hull_buffer(self):
        """The scene's hull records, concatenated ((words, 4) uint32; empty without hulls)."""
        if not self._hulls:
            return np.zeros((0, 4), np.uint32)
        out = np.zeros((self._hull_words, 4), np.uint32)
        for shape, at in self._hulls.values():
            out[at:at + shape.words] = shape.record(at)
        return out

    def addHeightfield(self, **options):
        """Static solid ground from a height grid (scenes.js:196-198; terrain.js cookHeightfield)."""
        from .terrain import cookHeightfield
        return self.addTerrain(cookHeightfield(**{k: v for k, v in options.items() if k in _HEIGHTFIELD_KEYS}),
                               **{k: v for k, v in options.items() if k not in _HEIGHTFIELD_KEYS})

    def addStaticMesh(self, **options):
        """Static triangle surfaces as thin solid prisms (scenes.js:199-201; terrain.js cookStaticMesh)."""
        from .terrain import cookStaticMesh
        return self.addTerrain(cookStaticMesh(**{k: v for k, v in options.items() if k in _MESH_KEYS}),
                               **{k: v for k, v in options.items() if k not in _MESH_KEYS})

    def addTerrain(self, cooked, **options):
        """Register cooked terrain pieces as static hull bodies (scenes.js:203-231; terrain.js createTerrainBodies)."""
        from .terrain import createTerrainBodies
        self.assertEditable()
        bodies = createTerrainBodies(self, cooked, options)
        return dict(bodies=bodies, triangleCount=cooked["triangleCount"], kind=cooked["kind"])

    def addBoxes(self, sizes, positions, rotations=None, density=1.0, friction=0.6, velocities=None,
                 angularVelocities=None):
        """Many boxes at once (a bulk form of addBox for large scenes); returns their first handle."""
        self.assertEditable()
        sizes = np.asarray(sizes, np.float64)
        positions = np.asarray(positions, np.float64)
        n = len(sizes)
        if sizes.shape != (n, 3) or positions.shape != (n, 3) or not (sizes > 0).all():
            raise ValueError("sizes and positions need n positive/finite triples")
        density = np.broadcast_to(np.asarray(density, np.float64), (n,))
        friction = np.broadcast_to(np.asarray(friction, np.float64), (n,))
        sx, sy, sz = sizes[:, 0], sizes[:, 1], sizes[:, 2]
        m = sx * sy * sz * density
        rec = np.zeros((n, 40), np.float32)
        rec[:, 0:3] = positions
        rec[:, 3] = friction
        if rotations is None:
            rec[:, 7] = 1.0
        else:
            q = np.asarray(rotations, np.float64)
            rec[:, 4:8] = q / np.linalg.norm(q, axis=1, keepdims=True)
        rec[:, 15] = 1.0
        rec[:, 16:19] = sizes
        rec[:, 19] = m
        rec[:, 20] = ((sy * sy + sz * sz) / 12) * m
        rec[:, 21] = ((sx * sx + sz * sz) / 12) * m
        rec[:, 22] = ((sx * sx + sy * sy) / 12) * m
        hx, hy, hz = sx * 0.5, sy * 0.5, sz * 0.5
        rec[:, 23] = np.sqrt(hx * hx + hy * hy + hz * hz)
        rec[:, 31] = 1.0
        if velocities is not None:
            rec[:, 32:35] = velocities
        if angularVelocities is not None:
            rec[:, 36:39] = angularVelocities
        rec[:, 39] = SHAPE_BOX
        first = self._count
        self._starts.append(self._count)
        self._chunks.append(rec)
        self._count += n
        return first

    def record(self, body):
        """The 40-float record of a built body (a copy)."""
        if not (0 <= body < self._count):
            raise IndexError("no such body")
        k = bisect.bisect_right(self._starts, body) - 1
        return self._chunks[k][body - self._starts[k]].copy()

    # --- constraints -----------------------------------------------------------------------------------------------

    def _check(self, a, b):
        if b is None or a == b or not (0 <= b < self._count) or (a is not None and not (0 <= a < self._count)):
            raise ValueError("Constraint endpoints must be distinct bodies in this scene")

    def addJoint(self, a, b, anchorA=(0, 0, 0), anchorB=(0, 0, 0), linearStiffness=math.inf, angularStiffness=0.0,
                 breakForce=math.inf, breakOnPull=False):
        """Ball joint plus optional angle lock (scenes.js:508-540; forces.ts:68-110). a = None: anchorA is a world point."""
        self.assertEditable()
        self._check(a, b)
        for v in (linearStiffness, angularStiffness, breakForce):
            if v != math.inf:
                _positive(v, "joint stiffness/breakForce", True)
        rA = _vector(anchorA, 3, "anchorA")
        rB = _vector(anchorB, 3, "anchorB")
        frac = -breakForce if (breakOnPull and breakForce != math.inf) else breakForce
        sA = self.record(a)[16:19] if a is not None else np.zeros(3, np.float32)
        sB = self.record(b)[16:19]
        s = [float(sA[k]) + float(sB[k]) for k in range(3)]
        torque_arm = s[0] * s[0] + s[1] * s[1] + s[2] * s[2]
        rec = np.zeros(32, np.float32)
        rec[3] = _finite(linearStiffness)
        rec[7] = _finite(angularStiffness)
        rec[11] = _finite(frac)
        rec[15] = torque_arm
        rec[24:27] = rA
        rec[28:31] = rB
        self._joints.append((rec, (T_JOINT, -1 if a is None else int(a), int(b), 0)))
        return len(self._joints) - 1

    def addSpring(self, a, b, anchorA=(0, 0, 0), anchorB=(0, 0, 0), stiffness=1000.0, rest=-1.0,
                  breakStrain=math.inf, breakForce=math.inf):
        """Spring between body-local anchors; rest -1 takes their initial separation. A finite breakStrain (extension
        over the rest length) or breakForce (tension) lets it tear for good (scenes.js:546-575; springMaterial.js)."""
        self.assertEditable()
        self._check(a, b)
        if a is None:
            raise ValueError("A spring requires two bodies")
        _positive(stiffness, "stiffness", True)
        if rest != -1:
            _positive(rest, "rest", True)
        rA = _vector(anchorA, 3, "anchorA")
        rB = _vector(anchorB, 3, "anchorB")
        if rest < 0:
            ra, rb = self.record(a), self.record(b)
            pa = np.add(_rotate([float(x) for x in ra[4:8]], rA), ra[0:3].astype(np.float64))
            pb = np.add(_rotate([float(x) for x in rb[4:8]], rB), rb[0:3].astype(np.float64))
            d = pa - pb
            rest = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        strain, force = _break_limits(breakStrain, breakForce)
        rec = np.zeros(32, np.float32)
        rec[3] = float(stiffness)
        rec[11] = strain
        rec[15] = force
        rec[24:27] = rA
        rec[27] = rest
        rec[28:31] = rB
        self._joints.append((rec, (T_SPRING, int(a), int(b), 0)))
        return len(self._joints) - 1

    def addMotor(self, a, b, axisA=(0, 1, 0), axisB=(0, 1, 0), speed=1.0, maxTorque=100.0):
        """An angular motor driving B's twist about the hinge axis (body-local axisA in A, axisB in B) at `speed`
        rad/s with at most `maxTorque` (scenes.js:572-582; angularConstraints.js). a = None: the world."""
        self.assertEditable()
        self._check(a, b)
        rec = _angular_record("motor", axisA, axisB, speed=speed, maxTorque=maxTorque)
        self._joints.append((rec, (T_MOTOR, -1 if a is None else int(a), int(b), 0)))
        return len(self._joints) - 1

    def addHinge(self, a, b, anchorA=(0, 0, 0), anchorB=(0, 0, 0), axisA=(0, 1, 0), axisB=(0, 1, 0), span=1.0,
                 minAngle=None, maxAngle=None, motor=None):
        """Two ball joints `span` apart along the axis, plus an angle limit when minAngle or maxAngle is given and a
        motor when `motor` (speed, maxTorque) is (scenes.js:587-630). Returns {joints, limit, motor} slots."""
        self.assertEditable()
        self._check(a, b)
        _angular_record("limit", axisA, axisB, minAngle=-math.pi if minAngle is None else minAngle,
                        maxAngle=math.pi if maxAngle is None else maxAngle)  # validates
        span = _positive(span, "span")
        ua, ub = _unit(axisA), _unit(axisB)
        aa, bb = _vector(anchorA, 3, "anchorA"), _vector(anchorB, 3, "anchorB")
        joints = [self.addJoint(a, b, [aa[i] + sign * span * 0.5 * ua[i] for i in range(3)],
                                [bb[i] + sign * span * 0.5 * ub[i] for i in range(3)]) for sign in (-1, 1)]
        limit = motor_slot = None
        if minAngle is not None or maxAngle is not None:
            rec = _angular_record("limit", axisA, axisB, minAngle=-math.pi if minAngle is None else minAngle,
                                  maxAngle=math.pi if maxAngle is None else maxAngle)
            self._joints.append((rec, (T_LIMIT, -1 if a is None else int(a), int(b), 0)))
            limit = len(self._joints) - 1
        if motor:
            motor_slot = self.addMotor(a, b, axisA, axisB, **motor)
        return {"joints": joints, "limit": limit, "motor": motor_slot}

    def ignoreCollision(self, a, b):
        """The pair never collides (ref/forces.ts IgnoreCollision)."""
        self.assertEditable()
        self._check(a, b)
        self._ignore.append((int(a), int(b)))

    # --- packing ---------------------------------------------------------------------------------------------------

    def packed(self):
        """(bodies (n, 40) f32, joints (J, 32) f32, info (J, 4) i32, ignore (K, 2) u32, filters (n, 2) u32 or None)."""
        bodies = np.ascontiguousarray(np.concatenate(self._chunks) if self._chunks else np.zeros((0, 40), np.float32))
        joints = np.ascontiguousarray(np.array([r for r, _ in self._joints], np.float32).reshape(-1, 32))
        info = np.ascontiguousarray(np.array([i for _, i in self._joints], np.int32).reshape(-1, 4))
        ignore = np.ascontiguousarray(np.array(self._ignore, np.uint32).reshape(-1, 2))
        filters = None
        if self._filters:
            filters = np.full((len(bodies), 2), 0xFFFFFFFF, np.uint32)
            for i, (g, m) in self._filters.items():
                filters[i] = (g, m)
        return bodies, joints, info, ignore, filters

    def world_params(self):
        """The 16-float per-world parameter block the CPU restatement takes (dt, gravity, betas, gamma, alpha, up,
        flags and step as bits)."""
        p = np.zeros(16, np.float32)
        p[0] = self.params["dt"]
        p[1] = self.params["gravity"]
        p[2] = self.params["betaLin"]
        p[3] = self.params["betaAng"]
        p[4] = self.params["gamma"]
        p[5] = self.params["alpha"]
        p[6:9] = self.up
        p.view(np.uint32)[9] = self.params["flags"]
        p.view(np.uint32)[10] = 0
        return p

"""Body edits staged on the host before they reach the GPU.

A command is the 12-word record of the reference's CommandStaging (src/gpu/bodyCommands2D.js:16-58):
u32 body, u32 kind, u32 has point, 0 | f32 value xyzw | f32 point xyzw. Kinds follow its editKeys
(src/native/runtime.js:36-44): 0 linearVelocity, 1 angularVelocity, 2 impulse, 3 force (value.w = dt), 4 teleport
(position; rotation in the point words), 5 angularImpulse, 6 torque (value.w = dt). Validation and ordering follow
runtime.js:16-175 (components, editBodies, stageEdits), solver3D.js:130-206 (the single-body setters, teleport) and
packedPoses.js:3-72 (setPoses): a whole batch is checked before anything is queued.
"""
import math

import numpy as np

EDIT_KEYS = ("linearVelocity", "angularVelocity", "impulse", "force", "", "angularImpulse", "torque")
LINEAR_VELOCITY, ANGULAR_VELOCITY, IMPULSE, FORCE, TELEPORT, ANGULAR_IMPULSE, TORQUE = range(7)
_UNSUPPORTED = ("isTrigger",)


def components(value, count, name):
    """runtime.js:16-32: an {x, y, z[, w]} object or a sequence of exactly `count` finite numbers."""
    keys = "xyzw"[:count]
    if isinstance(value, dict) and "x" in value:
        a = [value.get(k) for k in keys]
    elif hasattr(value, "x") and not isinstance(value, (list, tuple, np.ndarray)):
        a = [getattr(value, k, None) for k in keys]
    else:
        a = list(value) if value is not None else []
    try:
        ok = len(a) == count and all(v is not None and math.isfinite(float(v)) for v in a)
    except (TypeError, ValueError):
        ok = False
    if not ok:
        raise ValueError(f"{name} requires {count} finite components")
    return [float(v) for v in a]


def unit_rotation(rotation):
    """teleport's normalisation (solver3D.js:177-186): q / hypot(q) in double precision."""
    q = components(rotation, 4, "rotation")
    length = math.hypot(*q)
    if not length:
        raise ValueError("rotation must be nonzero")
    return [v / length for v in q]


def coefficient(value):
    """A restitution coefficient: a finite number in [0, 1] (appGpuSolver3D.js:455-467)."""
    try:
        v = float(value)
    except (TypeError, ValueError):
        v = math.nan
    if not (math.isfinite(v) and 0 <= v <= 1):
        raise ValueError("restitution must be between 0 and 1")
    return v


def mask(value, name="Collision masks"):
    if isinstance(value, bool) or not isinstance(value, (int, np.integer)) or value < 0 or value > 0xFFFFFFFF:
        raise ValueError(f"{name} must be unsigned 32-bit integers")
    return int(value)


class CommandStaging:
    """Growable command records (n x 12 u32, float payloads stored as f32 bits)."""

    def __init__(self):
        self.words = np.zeros((64, 12), np.uint32)
        self.count = 0

    def reserve(self, count):
        if count <= len(self.words):
            return
        grown = np.zeros((max(count, 2 * len(self.words)), 12), np.uint32)
        grown[:self.count] = self.words[:self.count]
        self.words = grown

    def append(self, index, kind, value, point=None, has_point=False):
        self.reserve(self.count + 1)
        row = self.words[self.count]
        row[:] = 0
        row[0] = index
        row[1] = kind
        row[2] = 1 if has_point else 0
        f = row.view(np.float32)
        for k, v in enumerate(value[:4]):
            f[4 + k] = v
        if point is not None:
            for k, v in enumerate(point[:4]):
                f[8 + k] = v
        self.count += 1

    def set_float(self, word, value):
        """Overwrite one float word of the last record (dt in value.w for forces and torques)."""
        self.words[self.count - 1].view(np.float32)[word] = value

    def extend(self, other):
        self.reserve(self.count + other.count)
        self.words[self.count:self.count + other.count] = other.words[:other.count]
        self.count += other.count

    def records(self):
        return self.words[:self.count]

    def clear(self):
        self.count = 0


def stage_edits(edits, body_index, dt_of, commands):
    """stageEdits (runtime.js:60-175) into `commands` (cleared first): per edit, a teleport first, then the remaining
    kinds in editKeys order. Returns the property edits (body, group, collidesWith, restitution, allowSleep) to apply
    afterwards (group None: no filter change; restitution or allowSleep None: no change).

    body_index(body) -> live slot (raises otherwise); dt_of(slot) -> the body's world time step."""
    commands.clear()
    properties = []
    for edit in edits:
        index = body_index(edit.get("body"))
        unsupported = [k for k in _UNSUPPORTED if edit.get(k) is not None]
        if unsupported:
            raise NotImplementedError(f"{unsupported[0]} edits are not available on the hip backend")
        group, collides = edit.get("group"), edit.get("collidesWith")
        restitution = edit.get("restitution")
        if restitution is not None:
            restitution = coefficient(restitution)
        allow = edit.get("allowSleep")
        if allow is not None and not isinstance(allow, bool):
            raise ValueError("allowSleep must be a boolean")
        if group is not None or collides is not None or restitution is not None or allow is not None:
            if collides is not None and group is None:
                raise ValueError("A collision-mask edit requires group as well")
            properties.append((index, None if group is None else mask(group),
                               mask(collides) if collides is not None else 0xFFFFFFFF, restitution, allow))
        point = edit.get("worldPoint")
        point = None if point is None else components(point, 3, "worldPoint")
        if edit.get("position") is not None:
            p = components(edit["position"], 3, "position")
            q = unit_rotation(edit["rotation"]) if edit.get("rotation") is not None else None
            commands.append(index, TELEPORT, p, q, q is not None)
        elif edit.get("rotation") is not None:
            raise ValueError("A rotation edit requires position; use teleport for a full pose")
        for kind, key in enumerate(EDIT_KEYS):
            if kind == TELEPORT or edit.get(key) is None:
                continue
            v = components(edit[key], 3, key)
            with_point = kind in (IMPULSE, FORCE) and point is not None
            commands.append(index, kind, v, point if with_point else None, with_point)
            if kind in (FORCE, TORQUE):
                commands.set_float(7, dt_of(index))
    return properties


def stage_poses(indices, positions, rotations, body_count, commands):
    """setPoses (solver3D.js:188-206; packedPoses.js:3-72): unique ascending slots, xyz positions, optional xyzw
    rotations scaled by 1 / hypot (double precision)."""
    if not (isinstance(indices, np.ndarray) and indices.dtype == np.uint32 and isinstance(positions, np.ndarray)
            and positions.dtype == np.float32 and (rotations is None or (isinstance(rotations, np.ndarray)
                                                                         and rotations.dtype == np.float32))):
        raise TypeError("Pose batches require Uint32 indices and Float32 pose arrays")
    n = len(indices)
    if positions.size != 3 * n or (rotations is not None and rotations.size != 4 * n):
        raise ValueError("Provide xyz positions and optional xyzw rotations per body")
    previous = -1
    for i in indices:
        if i >= body_count:
            raise ValueError("A live 3D GPU body index is required")
        if int(i) <= previous:
            raise ValueError("Pose batch slots must be unique and ascending")
        previous = int(i)
    pos = positions.reshape(-1, 3).astype(np.float64)
    if not np.isfinite(pos).all():
        raise ValueError("Pose positions must be finite")
    rot = None
    if rotations is not None:
        rot = rotations.reshape(-1, 4).astype(np.float64)
        length = np.array([math.hypot(*r) for r in rot])
        if not (np.isfinite(length).all() and (length > 0).all()):
            raise ValueError("Pose rotations must be finite and nonzero")
        rot = rot * (1.0 / length)[:, None]
    commands.reserve(commands.count + n)
    for k in range(n):
        commands.append(int(indices[k]), TELEPORT, list(pos[k]), list(rot[k]) if rot is not None else [0, 0, 0, 1],
                        rot is not None)

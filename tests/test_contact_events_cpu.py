"""The oracle's contact events (oracle/contact_events.py) on scripted scenes, pinned by what the reference's
ContactWatch states: a falling sphere begins touching the floor in the step its first manifold appears, with the
normal from the floor to the sphere, the point at the sphere's bottom and an impulse of about the momentum it loses;
a teleport away ends the pair in the next step; a reset lifetime begins the pair again without an end; unwatched
pairs report nothing."""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
from contact_events import BEGIN, END, ContactEvents  # noqa: E402
from oracle_lib import OracleWorld  # noqa: E402


def _scene(avbd):
    s = avbd.AvbdScene3D(gravity=-9.81)
    s.addBox([20, 1, 20], density=0, position=[0, -0.5, 0])
    s.addSphere(0.5, position=[0, 0.9, 0], velocity=[0, -2, 0])
    s.addBox([1, 1, 1], position=[5, 0.5, 0])
    return s


def _run(world, ev, steps, counter):
    for _ in range(steps):
        world.step()
        counter[0] += 1
        ev.after_step(world.manifolds(), world.contact_words(), world.bodies(), counter[0], 1 / 60)


def test_begin_end_and_reset(avbd):
    w = OracleWorld([_scene(avbd)], precision=32)
    ev, n = ContactEvents(3), [0]
    _run(w, ev, 1, n)
    first = ev.read()
    assert {(e[0], e[1], e[2]) for e in first} == {(BEGIN, 2, 0)}, first  # the resting box begins at once
    steps = 0
    while not any(e[1] == 1 for e in ev.log):
        _run(w, ev, 1, n)
        steps += 1
        assert steps < 60, "the sphere never landed"
    (kind, a, b, step, point, impulse, normal), = [e for e in ev.read() if e[1] == 1]
    assert kind == BEGIN and b == 0 and step == n[0]
    assert np.allclose(normal, [0, 1, 0], atol=1e-5) and abs(point[1]) < 0.05 and abs(point[0]) < 1e-4
    mass = 4 / 3 * np.pi * 0.5 ** 3
    assert 0.2 * mass * 2 < impulse < 3 * mass * 3, impulse  # about the momentum the landing takes away
    _run(w, ev, 5, n)
    assert ev.read() == []  # resting: no events
    cmd = np.zeros(12, np.uint32)
    cmd[0], cmd[1] = 1, 4
    cmd.view(np.float32)[4:7] = [0, 5, 0]
    w.apply_commands(cmd[None])
    _run(w, ev, 1, n)
    assert [(e[0], e[1], e[2]) for e in ev.read()] == [(END, 1, 0)]
    ev.reset([2])
    _run(w, ev, 1, n)
    assert [(e[0], e[1], e[2]) for e in ev.read()] == [(BEGIN, 2, 0)]  # a new lifetime: begins again, no end
    ev.set_watched([0, 2], [0, 0])
    w.apply_commands(np.concatenate([cmd[None], cmd[None]]))
    c2 = cmd.copy()
    c2[0] = 2
    c2.view(np.float32)[4:7] = [10, 5, 0]  # the box leaves the floor, away from the sphere
    w.apply_commands(c2[None])
    _run(w, ev, 2, n)
    assert ev.read() == []  # neither body of the pair is watched

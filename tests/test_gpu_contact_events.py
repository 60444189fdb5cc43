"""Contact events (contacts.ts ContactWatch) on the GPU against the oracle's restatement (oracle/contact_events.py):
the same begin and end events, step by step, with the same points, impulses and normals bit for bit, for a pile of
boxes and spheres landing and settling, with a teleport, a reset lifetime and an unwatched subset."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
from contact_events import ContactEvents  # noqa: E402
from oracle_lib import OracleWorld  # noqa: E402

pytestmark = pytest.mark.gpu


def scene(avbd):
    rng = np.random.default_rng(21)
    s = avbd.AvbdScene3D(gravity=-9.81)
    s.addBox([30, 1, 30], density=0, position=[0, -0.5, 0])
    for k in range(60):
        pos = [float(rng.uniform(-3, 3)), float(0.8 + 0.5 * k), float(rng.uniform(-3, 3))]
        if k % 2:
            s.addSphere(float(rng.uniform(0.3, 0.5)), position=pos)
        else:
            s.addBox(list(rng.uniform(0.5, 1.0, 3)), position=pos, rotation=list(rng.normal(size=4)))
    return s


def _key(e):
    return (e[3], e[0], e[1], e[2], e[4], e[5], e[6])


def test_events_match_the_oracle(avbd):
    world = scene(avbd).createSolver(avbd.createSolverPipeline("hip"), {"colorCap": 32, "colorRounds": 16})
    oracle = OracleWorld([scene(avbd)], precision=32, color_cap=32, color_rounds=16)
    world.watchContacts(maxPairs=4096, maxEvents=8192)
    ev = ContactEvents(world.bodyCount)
    total = 0
    for step in range(1, 161):
        if step == 90:
            cmd = np.zeros(12, np.uint32)
            cmd[0], cmd[1] = 5, 4
            cmd.view(np.float32)[4:7] = [12, 3, 12]
            world.teleport(5, [12, 3, 12])
            oracle.apply_commands(cmd[None])
        if step == 100:
            world.resetContactHistory([3, 7])
            ev.reset([3, 7])
        if step == 110:
            world.watchContacts(indices=list(range(0, world.bodyCount, 2)))
            ev.set_watched(range(world.bodyCount), [int(i % 2 == 0) for i in range(world.bodyCount)])
        world.step()
        oracle.step()
        ev.after_step(oracle.manifolds(), oracle.contact_words(), oracle.bodies(), step, 1 / 60)
        got = world.readContactEvents()
        assert got["dropped"] == 0 and not got["full"]
        g = sorted((e["kind"], e["a"], e["b"], e["step"], tuple(e["point"]), e["impulse"], tuple(e["normal"]))
                   for e in got["events"])
        want = sorted(ev.read())
        assert [_key(e) for e in sorted(g, key=_key)] == [_key(e) for e in sorted(want, key=_key)], f"step {step}"
        total += len(g)
    assert total > 60
    world.destroy()

"""Pins of the oracle's GPU-only warm-start rules by checks the rules' definitions fix (CPU only), on scripted scenes:

* contactPenaltyMin (reference/three-avbd/src/avbd3d/gpu/wgsl-collision.ts:631-636, 674, 759): a new contact's normal
  penalty starts at clamp(min(mA, mB) / dt², 1, 1e10) with a static body counting as infinitely heavy, and at 1 with
  the flag off; with no gravity and a 5 mm overlap (inside the 0.01 m margin) no row is active, so the value survives a
  whole step unchanged;
* matchNearest (wgsl-collision.ts:768-795): a point whose feature key is new takes the multipliers of the previous
  point nearest in A-local anchor within 0.05 × the smaller body's smallest side (the later one on a tie), scaled by
  αγ; none when no previous point is that near or with the flag off — checked against a brute-force search over
  crafted previous points;
* reuseContacts (wgsl-collision.ts:20-47, 638-678): a pair whose bodies stayed within 2 mm and 0.002 rad of their
  reference poses keeps its points (anchors bit for bit and the generation step of its manifold word); a move of 1 cm
  or 0.005 rad generates them again at the current step; with the flag off every step generates them;
* the 1e-5 relative pivot floor (reference/three-avbd/src/avbd3d/gpu/layout.ts:105-106, wgsl-solve.ts:416-470): for
  A = L D Lᵀ with a unit lower L and one pivot d_k below the floor, the solve equals numpy's exact solution of
  (A + (1e-5 A_kk − d_k) e_k e_kᵀ) x = b: raising a pivot during the factorisation factorises A with that diagonal
  entry raised by the same amount (the modified-Cholesky identity);
* start-at-rest (wgsl-solve.ts:259-300): a body slower than 0.1 (|v| + |ω| r) starts its solve from its pose, a faster
  one from x + v dt (+ the adaptive gravity term); the consequence on a resting stack is less creep per step.
"""
import ctypes
import importlib

import numpy as np
import pytest

from oracle_lib import OracleWorld, lib

avbd = importlib.import_module("babylon-avbd_amd")
FLAGS = avbd._native.FLAGS_DEFAULT
MATCH_NEAREST, REUSE, START_AT_REST, MASS_PENALTY = 1, 4, 8, 16
f32 = np.float32


def _contacts(o):
    """(manifold rows, contact words, contact floats) of the oracle's last collision."""
    mf = o.manifolds()
    ct = o.contact_words()
    return mf, ct, ct.view(np.float32)


def _penalty_min(m_a, m_b, dt):
    big = f32(3.4e38)
    ma = big if m_a <= 0 else f32(m_a)
    mb = big if m_b <= 0 else f32(m_b)
    return min(max(min(ma, mb) / (f32(dt) * f32(dt)), f32(1)), f32(1e10))


@pytest.mark.parametrize("density", [1e-4, 1.0, 7.8, 1e9])
@pytest.mark.parametrize("flag", [True, False])
def test_contact_penalty_min(density, flag):
    flags = FLAGS if flag else FLAGS & ~MASS_PENALTY
    s = avbd.AvbdScene3D(gravity=0.0, flags=flags)
    s.addBox([20, 1, 20], density=0, position=[0, -0.5, 0])
    box = s.addBox([1, 1, 1], density=density, position=[0, 0.5 - 0.005, 0])
    mass = float(s.record(box)[19])
    o = OracleWorld([s], precision=32)
    o.step(1)
    mf, ct, f = _contacts(o)
    assert len(mf) == 1 and (mf[0, 3] & 15) == 4
    want = _penalty_min(mass, 0.0, s.params["dt"]) if flag else f32(1)
    assert np.all(f[:, 8] == want), (f[:, 8], want)
    assert np.all(f[:, 9:11] == f32(1)), "tangent penalties start at the minimum"
    assert np.all(f[:, 12:15] == 0), "no row was active: no force"


def test_contact_penalty_min_between_two_dynamic_bodies():
    s = avbd.AvbdScene3D(gravity=0.0)
    a = s.addBox([1, 1, 1], density=2.0, position=[0, 0, 0])
    b = s.addBox([1, 1, 1], density=5.0, position=[1 - 0.005, 0, 0])
    o = OracleWorld([s], precision=32)
    o.step(1)
    _, _, f = _contacts(o)
    assert len(f) == 4
    assert np.all(f[:, 8] == _penalty_min(s.record(a)[19], s.record(b)[19], s.params["dt"]))


def _resting_pair(flags):
    """A unit box overlapping a static ground by 5 mm, no gravity: its contacts do not change from step to step."""
    s = avbd.AvbdScene3D(gravity=0.0, flags=flags)
    s.addBox([20, 1, 20], density=0, position=[0, -0.5, 0])
    s.addBox([1, 1, 1], position=[0.13, 0.5 - 0.005, -0.21], rotation=[0, 0.2, 0, 0.98])
    return s


def _records(o):
    """The oracle's bodies as 40-float records, the moved-step stamp as u32 bits in inertialPos.w (as a GPU world
    reads them back)."""
    b64 = o.bodies()
    b = b64.astype(np.float32)
    b.view(np.uint32)[:, 27] = b64[:, 27].astype(np.uint32)
    return b


def _state_of(o, s):
    mf, ct, _ = _contacts(o)
    return dict(bodies=_records(o), joints=o.joints().astype(np.float32), refs=o.refs(),
                colors=o.colors(), manifolds=mf, contacts=ct, steps=np.array([1], np.uint32))


@pytest.mark.parametrize("nearest", [True, False])
def test_match_nearest_against_brute_force(nearest):
    flags = (FLAGS & ~REUSE) if nearest else (FLAGS & ~REUSE & ~MATCH_NEAREST)
    s = _resting_pair(flags)
    o = OracleWorld([s], precision=32)
    o.partial_step(1)
    state = _state_of(o, s)
    new_mf, new_ct, new_f = _contacts(o)  # the points the next collision finds again (nothing moves)
    # Craft the previous points: feature keys that never match, anchors offset from the new points' anchors by
    # 0.01 m, 0.02 m (inside 0.05 x smallest side = 0.05 m), 0.08 m and 0.3 m (outside), distinct multipliers
    rng = np.random.default_rng(4)
    prev = new_ct.copy()
    pf = prev.view(np.float32)
    offsets = [0.01, 0.02, 0.08, 0.3]
    order = rng.permutation(4)
    for c in range(4):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        pf[c, 0:3] = new_f[order[c], 0:3] + f32(offsets[c]) * d.astype(np.float32)
        pf[c, 12:15] = [-3.0 - c, 0.5 + c, -0.25 * c]
        pf[c, 8:11] = [1000.0 * (c + 1), 200.0 + c, 300.0 + c]
    prev[:, 3] = 0x00ABCD00 + np.arange(4, dtype=np.uint32)  # no STICK_BIT: anchors are not carried
    state["contacts"] = prev
    o.seed(state)
    o.partial_step(1)
    mf, ct, f = _contacts(o)
    assert np.array_equal(ct[:, 3] & 0x7FFFFFFF, new_ct[:, 3] & 0x7FFFFFFF), "the same points are found"
    p = s.params
    pmin = _penalty_min(s.record(1)[19], 0.0, p["dt"])
    for i in range(4):
        dist = np.linalg.norm(pf[:, 0:3].astype(np.float64) - f[i, 0:3].astype(np.float64), axis=1)
        inside = np.flatnonzero(dist <= 0.05)
        if nearest and len(inside):
            j = inside[np.argmin(dist[inside])]
            want_lam = pf[j, 12:15] * f32(p["alpha"]) * f32(p["gamma"])
            want_pen = np.minimum(np.maximum(pf[j, 8:11] * f32(p["gamma"]), f32(1)), f32(1e10))
            want_pen[0] = max(want_pen[0], pmin)
        else:
            want_lam = np.zeros(3, np.float32)
            want_pen = np.array([pmin, 1, 1], np.float32)
        assert np.array_equal(f[i, 12:15], want_lam), (i, f[i, 12:15], want_lam)
        assert np.array_equal(f[i, 8:11], want_pen), (i, f[i, 8:11], want_pen)
    if nearest:
        assert sum(np.linalg.norm(pf[:, 0:3] - f[i, 0:3], axis=1).min() <= 0.05 for i in range(4)) == 2


def _settled_box(flags):
    s = avbd.AvbdScene3D(gravity=-9.81, flags=flags)
    s.addBox([20, 1, 20], density=0, friction=0.8, position=[0, -0.5, 0])
    s.addBox([1, 1, 1], friction=0.8, position=[0, 0.55, 0])
    o = OracleWorld([s], precision=32)
    o.step(150)
    return s, o


def _seed_moved(o, move=(0, 0, 0), turn=0.0):
    """Seed the oracle from its own state with the box moved / turned about up (its reference pose kept)."""
    mf, ct, _ = _contacts(o)
    b = _records(o)
    b[1, 0:3] += np.asarray(move, np.float32)
    if turn:
        q = b[1, 4:8].astype(np.float64)
        r = np.array([0, np.sin(turn / 2), 0, np.cos(turn / 2)])
        x1, y1, z1, w1 = r
        x2, y2, z2, w2 = q
        b[1, 4:8] = [w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2]
    steps = int(mf[0, 3] >> 4) if len(mf) else 0
    o.seed(dict(bodies=b, joints=o.joints().astype(np.float32), refs=o.refs(), colors=o.colors(), manifolds=mf,
                contacts=ct, steps=np.array([150], np.uint32)))
    return steps


def test_reuse_contacts_keeps_points_of_still_pairs():
    s, o = _settled_box(FLAGS)
    mf0, ct0, _ = _contacts(o)
    gen0 = int(mf0[0, 3] >> 4)
    for _ in range(20):  # at rest: the same points, generated at the same step
        o.step(1)
        mf, ct, _ = _contacts(o)
        assert int(mf[0, 3] >> 4) == gen0
        assert np.array_equal(ct[:, 0:3], ct0[:, 0:3]) and np.array_equal(ct[:, 4:7], ct0[:, 4:7])
    for move, turn, reused in (((0.001, 0, 0), 0.0, True), ((0, 0, 0), 0.001, True), ((0.01, 0, 0), 0.0, False),
                               ((0, 0, 0), 0.005, False)):
        s, o = _settled_box(FLAGS)
        mf0, ct0, _ = _contacts(o)
        _seed_moved(o, move, turn)
        o.partial_step(1)
        mf, ct, _ = _contacts(o)
        gen = int(mf[0, 3] >> 4)
        if reused:
            assert gen == int(mf0[0, 3] >> 4), (move, turn)
            assert np.array_equal(ct[:, 0:3], ct0[:, 0:3]) and np.array_equal(ct[:, 4:7], ct0[:, 4:7]), (move, turn)
        else:
            assert gen == 151, (move, turn, gen)


def test_reuse_off_generates_every_step():
    s, o = _settled_box(FLAGS & ~REUSE)
    for k in range(5):
        o.step(1)
        mf, _, _ = _contacts(o)
        assert int(mf[0, 3] >> 4) == 151 + k


def test_pivot_floor_is_the_exact_solve_of_the_floored_factorisation():
    rng = np.random.default_rng(2)
    D = ctypes.POINTER(ctypes.c_double)
    for trial in range(40):
        L = np.eye(6)
        L[np.tril_indices(6, -1)] = rng.uniform(0.3, 1.0, 15) * rng.choice([-1.0, 1.0], 15)
        d = rng.uniform(0.5, 2.0, 6)
        k = int(rng.integers(1, 6))
        above = float((L[k, :k] ** 2 * d[:k]).sum())
        d[k] = rng.uniform(-1e-7, 1e-7) * above  # far below 1e-5 x its diagonal (above + d_k)
        A = L @ np.diag(d) @ L.T
        b = rng.normal(size=6)
        floor = 1e-5 * A[k, k]
        # raising pivot k from d_k to the floor is factorising A with d_k's shortfall added to A_kk
        bumped = A.copy()
        bumped[k, k] += floor - d[k]
        want = np.linalg.solve(bumped, b)
        full = np.ascontiguousarray(A)
        x = np.zeros(6)
        lib().gs_probe_solve6(full.ctypes.data_as(D), b.ctypes.data_as(D), 1e-5, x.ctypes.data_as(D))
        assert np.allclose(x, want, rtol=1e-6, atol=1e-9 * np.abs(want).max()), (trial, k, x, want)
        # and a well-conditioned system is solved exactly
        A = np.ascontiguousarray(L @ np.diag(rng.uniform(0.5, 2.0, 6)) @ L.T)
        lib().gs_probe_solve6(A.ctypes.data_as(D), b.ctypes.data_as(D), 1e-5, x.ctypes.data_as(D))
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)


def test_start_at_rest_initial_guess():
    """Stage 2 (after the warm starts): the body's pose record holds its initial guess."""
    for flag in (True, False):
        flags = FLAGS if flag else FLAGS & ~START_AT_REST
        s = avbd.AvbdScene3D(gravity=-9.81, flags=flags)
        s.addBox([20, 1, 20], density=0, position=[0, -0.5, 0])
        slow = s.addBox([1, 1, 1], position=[-3, 2, 0], velocity=[0.06, 0, 0], angularVelocity=[0, 0.03, 0])
        fast = s.addBox([1, 1, 1], position=[3, 2, 0], velocity=[0.12, 0, 0])
        o = OracleWorld([s], precision=32)
        before = o.bodies()
        o.partial_step(2)
        b = o.bodies()
        dt = s.params["dt"]
        # the slow body: |v| + |ω| r = 0.06 + 0.03 x 0.866 < 0.1
        if flag:
            assert np.array_equal(b[slow, 0:8], before[slow, 0:8]), "a slow body starts from its pose"
        else:
            assert abs(b[slow, 0] - (before[slow, 0] + 0.06 * dt)) < 1e-6
        assert abs(b[fast, 0] - (before[fast, 0] + 0.12 * dt)) < 1e-6, "a fast body starts from x + v dt"
        # no vertical velocity and none last step: the adaptive gravity weight is 0
        assert abs(b[fast, 1] - before[fast, 1]) < 1e-7


def test_start_at_rest_reduces_creep_of_a_resting_stack():
    creep = {}
    for flag in (True, False):
        flags = FLAGS if flag else FLAGS & ~START_AT_REST
        s = avbd.AvbdScene3D(gravity=-9.81, flags=flags)
        s.addBox([20, 1, 20], density=0, friction=0.8, position=[0, -0.5, 0])
        for k in range(4):
            s.addBox([1, 1, 1], friction=0.8, position=[0, 0.5 + k, 0])
        o = OracleWorld([s], precision=32)
        o.step(200)
        prev = o.bodies()
        moves = []
        for _ in range(100):
            o.step(1)
            now = o.bodies()
            moves.append(np.abs(now[1:, 0:3] - prev[1:, 0:3]).max())
            prev = now
        creep[flag] = float(np.mean(moves))
    assert creep[True] < creep[False], creep

// ORACLE — test infrastructure only (see math.hpp).
//
// FP64 restatement of the upstream AVBD 3D demo as the reference vendors it (reference/three-avbd/src/avbd3d/ref/):
// body.ts (Rigid, constants), manifold.ts (Manifold), forces.ts (Joint, Spring, IgnoreCollision), solver.ts (Solver::
// step) and scenes.ts (the 14 demo scenes). The sequential order (bodies newest first, each body's forces in reverse
// creation order, duals in creation order) is kept, so the restatement reproduces the C++ demo's FP64 golden
// trajectories (reference/three-avbd/tests/fixtures/oracle3d/scene-*.json, test avbd3d-ref.test.ts:27-49).
#pragma once
#include <algorithm>
#include <limits>
#include <memory>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "collide.hpp"

namespace oracle {
namespace ref {

using R = double;
using V = V3<double>;
using Q = Q4<double>;
using M = M3<double>;

constexpr R PENALTY_MIN = 1.0;             // body.ts:8
constexpr R PENALTY_MAX = 10000000000.0;   // body.ts:9
constexpr R COLLISION_MARGIN = 0.01;       // body.ts:10
constexpr R STICK_THRESH = 0.00001;        // body.ts:11
constexpr R INF = std::numeric_limits<double>::infinity();

struct Solver;
struct Force;

/** body.ts:13-62. */
struct Rigid {
  std::vector<Force *> forces;
  V positionLin;
  Q positionAng;
  V initialLin;
  Q initialAng;
  V inertialLin;
  Q inertialAng;
  V velocityLin, velocityAng, prevVelocityLin;
  V size;
  R mass = 0;
  V moment;
  R friction = 0;
  R radius = 0;
  bool constrainedTo(const Rigid *other) const;
};

/** The per-body 6x6 system (body.ts:65-74). */
struct BodySystem {
  M lhsLin, lhsAng, lhsCross;
  V rhsLin, rhsAng;
};

struct Force {
  Solver *solver;
  Rigid *bodyA;  // null: world
  Rigid *bodyB;
  virtual ~Force() = default;
  virtual bool initialize() = 0;
  virtual void updatePrimal(Rigid *body, R alpha, BodySystem &sys) = 0;
  virtual void updateDual(R alpha) = 0;
  virtual int kind() const = 0;  // 0 manifold, 1 joint, 2 spring, 3 ignore
  void unlinkFromBodies() {
    for (Rigid *b : {bodyA, bodyB}) {
      if (!b) continue;
      auto it = std::find(b->forces.begin(), b->forces.end(), this);
      if (it != b->forces.end()) b->forces.erase(it);
    }
  }
};

inline bool Rigid::constrainedTo(const Rigid *other) const {
  for (const Force *f : forces)
    if ((f->bodyA == this && f->bodyB == other) || (f->bodyA == other && f->bodyB == this)) return true;
  return false;
}

struct Contact {
  uint32_t feature = 0;
  V rA, rB, C0, penalty, lambda;
  bool stick = false;
};

struct Solver {
  R dt = 1.0 / 60.0, gravity = -10, alpha = 0.99, betaLin = 10000, betaAng = 100, gamma = 0.999;
  int iterations = 10;
  // The package's CPU path (src/native/cpu/cpuSolver3D.ts:144-280, cpuPhysics3D.js:67-147, 304-363): gravity along
  // `up`, the contact mass-penalty floor, the face-biased box SAT and the hash-grid broadphase. Off: the upstream demo.
  bool package = false;
  V up = V(0, 0, 1);
  bool massPenalty = true, faceBias = true;
  // The colour-parallel variant of the package step (threads > 1): the same equations, each body's Newton step taken
  // with the bodies of its greedy colour in parallel (Gauss–Seidel by colour instead of by index), forces initialised
  // and duals updated in parallel. threads == 1: the reference's sequential order.
  int threads = 1;
  std::vector<std::unique_ptr<Rigid>> bodyStore;
  std::vector<Rigid *> bodies;
  std::vector<Force *> forces;                    // live, creation order
  std::vector<std::unique_ptr<Force>> forceStore;  // ownership

  Rigid *addBody(V size, R density, R friction, V position, V velocity = V()) {
    auto b = std::make_unique<Rigid>();
    b->size = size;
    b->friction = friction;
    b->positionLin = position;
    b->velocityLin = velocity;
    b->prevVelocityLin = velocity;
    R sx = size.x, sy = size.y, sz = size.z;
    b->mass = sx * sy * sz * density;
    b->moment = V(((sy * sy + sz * sz) / 12) * b->mass, ((sx * sx + sz * sz) / 12) * b->mass,
                  ((sx * sx + sy * sy) / 12) * b->mass);
    b->radius = length(V(sx * 0.5, sy * 0.5, sz * 0.5));
    bodies.push_back(b.get());
    bodyStore.push_back(std::move(b));
    return bodies.back();
  }
  template <class F, class... Args> F *addForce(Rigid *a, Rigid *b, Args &&...args) {
    auto f = std::make_unique<F>(std::forward<Args>(args)...);
    f->solver = this;
    f->bodyA = a;
    f->bodyB = b;
    forces.push_back(f.get());
    if (a) a->forces.push_back(f.get());
    b->forces.push_back(f.get());
    F *raw = f.get();
    forceStore.push_back(std::move(f));
    raw->created();
    return raw;
  }
  void step();
  void packageBroadphase();
  void packageStep();
  std::vector<uint32_t> colors;  // the colour-parallel variant's body colours of the current step
};

/** manifold.ts:67-218. */
struct Manifold : Force {
  std::vector<Contact> contacts;
  M basis;
  R friction = 0;
  void created() {}
  int kind() const override { return 0; }

  bool initialize() override {
    Rigid *A = bodyA, *B = bodyB;
    friction = std::sqrt(A->friction * B->friction);
    Box<R> bA = makeBox(A->positionLin, A->positionAng, A->size);
    Box<R> bB = makeBox(B->positionLin, B->positionAng, B->size);
    Sat<R> sat;
    Found<R> found = collideBoxes(bA, bB, solver->package && solver->faceBias, 16, sat);
    if (found.count) basis = orthonormal(neg(sat.n));
    std::vector<Contact> fresh;
    for (int i = 0; i < found.count; i++) {
      Contact c;
      c.feature = found.feature[i];
      V rA = rotateInv(A->positionAng, sub(found.xA[i], A->positionLin));
      V rB = rotateInv(B->positionAng, sub(found.xB[i], B->positionLin));
      const Contact *prev = nullptr;
      for (const Contact &o : contacts)
        if (o.feature == c.feature) {
          prev = &o;
          break;
        }
      if (!prev) {
        c.rA = rA;
        c.rB = rB;
      } else {
        c.rA = prev->stick ? prev->rA : rA;
        c.rB = prev->stick ? prev->rB : rB;
        c.penalty = prev->penalty;
        c.lambda = prev->lambda;
        c.stick = prev->stick;
      }
      fresh.push_back(c);
    }
    contacts = fresh;
    for (Contact &c : contacts) {
      V xA = transform(A->positionLin, A->positionAng, c.rA);
      V xB = transform(B->positionLin, B->positionAng, c.rB);
      c.C0 = mulv(basis, sub(xA, xB));
      c.C0.x += COLLISION_MARGIN;
      for (int i = 0; i < 3; i++) {
        c.lambda[i] = c.lambda[i] * solver->alpha * solver->gamma;
        c.penalty[i] = rclamp(c.penalty[i] * solver->gamma, PENALTY_MIN, PENALTY_MAX);
      }
      if (solver->package && solver->massPenalty) {  // CpuManifold3D.initialize (cpuPhysics3D.js:130-139)
        const R mass = std::min(A->mass > 0 ? A->mass : INF, B->mass > 0 ? B->mass : INF);
        c.penalty.x = std::max(c.penalty.x, rclamp(mass / (solver->dt * solver->dt), PENALTY_MIN, PENALTY_MAX));
      }
    }
    return !contacts.empty();
  }

  // evaluate() scratch (manifold.ts:111-144)
  V dqALin, dqAAng, dqBLin, dqBAng;
  M jBLin, jAAng, jBAng, K;
  V C, F;
  R frictionScale = 0, bounds = 0;

  static M angularJacobian(V r, const M &lin) {
    M o;
    for (int k = 0; k < 3; k++) {
      V rr = cross(r, row(lin, k));
      o.m[k * 3] = rr.x;
      o.m[k * 3 + 1] = rr.y;
      o.m[k * 3 + 2] = rr.z;
    }
    return o;
  }
  void computeDisplacements() {
    dqALin = sub(bodyA->positionLin, bodyA->initialLin);
    dqAAng = qsub(bodyA->positionAng, bodyA->initialAng);
    dqBLin = sub(bodyB->positionLin, bodyB->initialLin);
    dqBAng = qsub(bodyB->positionAng, bodyB->initialAng);
  }
  void evaluate(const Contact &c, R alpha) {
    V rAW = rotate(bodyA->positionAng, c.rA);
    V rBW = rotate(bodyB->positionAng, c.rB);
    jBLin = negm(basis);
    jAAng = angularJacobian(rAW, basis);
    jBAng = angularJacobian(rBW, jBLin);
    K = diag(c.penalty.x, c.penalty.y, c.penalty.z);
    C = scale(c.C0, 1 - alpha);
    C = addScaled(C, mulv(basis, dqALin), 1.0);
    C = addScaled(C, mulv(jBLin, dqBLin), 1.0);
    C = addScaled(C, mulv(jAAng, dqAAng), 1.0);
    C = addScaled(C, mulv(jBAng, dqBAng), 1.0);
    F = mulv(K, C);
    F = addScaled(F, c.lambda, 1.0);
    F.x = rmin(F.x, 0.0);
    bounds = std::fabs(F.x) * friction;
    frictionScale = std::sqrt(F.y * F.y + F.z * F.z);
    if (frictionScale > bounds && frictionScale > 0) {
      F.y *= bounds / frictionScale;
      F.z *= bounds / frictionScale;
    }
  }
  void updatePrimal(Rigid *body, R alpha, BodySystem &sys) override {
    computeDisplacements();
    bool isA = body == bodyA;
    for (const Contact &c : contacts) {
      evaluate(c, alpha);
      const M &jLin = isA ? basis : jBLin;
      const M &jAng = isA ? jAAng : jBAng;
      M jLinT = transpose(jLin), jAngT = transpose(jAng);
      M jAngTk = mul(jAngT, K);
      addm(sys.lhsLin, mul(mul(jLinT, K), jLin));
      addm(sys.lhsAng, mul(jAngTk, jAng));
      addm(sys.lhsCross, mul(jAngTk, jLin));
      sys.rhsLin = addScaled(sys.rhsLin, mulv(jLinT, F), 1.0);
      sys.rhsAng = addScaled(sys.rhsAng, mulv(jAngT, F), 1.0);
    }
  }
  void updateDual(R alpha) override {
    computeDisplacements();
    for (Contact &c : contacts) {
      evaluate(c, alpha);
      c.lambda = F;
      if (F.x < 0) c.penalty.x = rmin(c.penalty.x + solver->betaLin * std::fabs(C.x), PENALTY_MAX);
      if (frictionScale <= bounds) {
        c.penalty.y = rmin(c.penalty.y + solver->betaLin * std::fabs(C.y), PENALTY_MAX);
        c.penalty.z = rmin(c.penalty.z + solver->betaLin * std::fabs(C.z), PENALTY_MAX);
        c.stick = std::sqrt(C.y * C.y + C.z * C.z) < STICK_THRESH;
      }
    }
  }
};

inline M geometricStiffnessBallSocket(int k, V v) {  // forces.ts:55-61
  M o = diag(-v[k], -v[k], -v[k]);
  o.m[k] += v.x;
  o.m[3 + k] += v.y;
  o.m[6 + k] += v.z;
  return o;
}
inline V componentMin(V a, R b) { return V(rmin(a.x, b), rmin(a.y, b), rmin(a.z, b)); }

/** forces.ts:68-238. */
struct Joint : Force {
  V rA, rB, C0Lin, C0Ang, penaltyLin, penaltyAng, lambdaLin, lambdaAng;
  R stiffnessLin = INF, stiffnessAng = 0, fracture = INF, torqueArm = 0;
  bool broken = false;
  Joint(V a, V b, R kLin = INF, R kAng = 0, R frac = INF) : rA(a), rB(b), stiffnessLin(kLin), stiffnessAng(kAng), fracture(frac) {}
  int kind() const override { return 1; }
  void created() {
    V s = bodyA ? bodyA->size : V(0, 0, 0);
    torqueArm = lengthSq(V(s.x + bodyB->size.x, s.y + bodyB->size.y, s.z + bodyB->size.z));
  }
  V evaluateLin() const {
    V pA = bodyA ? transform(bodyA->positionLin, bodyA->positionAng, rA) : rA;
    return sub(pA, transform(bodyB->positionLin, bodyB->positionAng, rB));
  }
  V evaluateAng() const {
    V o = qsub(bodyA ? bodyA->positionAng : Q(), bodyB->positionAng);
    return scale(o, torqueArm);
  }
  bool initialize() override {
    C0Lin = evaluateLin();
    C0Ang = evaluateAng();
    for (int i = 0; i < 3; i++) {
      lambdaLin[i] = lambdaLin[i] * solver->alpha * solver->gamma;
      lambdaAng[i] = lambdaAng[i] * solver->alpha * solver->gamma;
      penaltyLin[i] = rclamp(penaltyLin[i] * solver->gamma, PENALTY_MIN, PENALTY_MAX);
      penaltyAng[i] = rclamp(penaltyAng[i] * solver->gamma, PENALTY_MIN, PENALTY_MAX);
    }
    penaltyLin = componentMin(penaltyLin, stiffnessLin);
    penaltyAng = componentMin(penaltyAng, stiffnessAng);
    return !broken;
  }
  void updatePrimal(Rigid *body, R alpha, BodySystem &sys) override {
    bool isA = body == bodyA;
    if (lengthSq(penaltyLin) > 0) {
      M K = diag(penaltyLin.x, penaltyLin.y, penaltyLin.z);
      V C = evaluateLin();
      if (stiffnessLin == INF) C = addScaled(C, C0Lin, -alpha);
      V F = mulv(K, C);
      F = addScaled(F, lambdaLin, 1.0);
      R sgn = isA ? 1 : -1;
      M jLin = diag(sgn, sgn, sgn);
      M jAng = isA ? skew(neg(rotate(bodyA->positionAng, rA))) : skew(rotate(bodyB->positionAng, rB));
      M jLinT = transpose(jLin), jAngT = transpose(jAng);
      M jAngTk = mul(jAngT, K);
      addm(sys.lhsLin, mul(mul(jLinT, K), jLin));
      addm(sys.lhsAng, mul(jAngTk, jAng));
      addm(sys.lhsCross, mul(jAngTk, jLin));
      V r = isA ? rotate(bodyA->positionAng, rA) : neg(rotate(bodyB->positionAng, rB));
      M H = scalem(geometricStiffnessBallSocket(0, r), F.x);
      addm(H, scalem(geometricStiffnessBallSocket(1, r), F.y));
      addm(H, scalem(geometricStiffnessBallSocket(2, r), F.z));
      addm(sys.lhsAng, diagonalize(H));
      sys.rhsLin = addScaled(sys.rhsLin, mulv(jLinT, F), 1.0);
      sys.rhsAng = addScaled(sys.rhsAng, mulv(jAngT, F), 1.0);
    }
    if (lengthSq(penaltyAng) > 0) {
      M K = diag(penaltyAng.x, penaltyAng.y, penaltyAng.z);
      V C = evaluateAng();
      if (stiffnessAng == INF) C = addScaled(C, C0Ang, -alpha);
      V F = mulv(K, C);
      F = addScaled(F, lambdaAng, 1.0);
      R s = (isA ? 1 : -1) * torqueArm;
      M jAng = diag(s, s, s);
      M jAngT = transpose(jAng);
      addm(sys.lhsAng, mul(mul(jAngT, K), jAng));
      sys.rhsAng = addScaled(sys.rhsAng, mulv(jAngT, F), 1.0);
    }
  }
  void updateDual(R alpha) override {
    if (lengthSq(penaltyLin) > 0) {
      M K = diag(penaltyLin.x, penaltyLin.y, penaltyLin.z);
      V C = evaluateLin();
      if (stiffnessLin == INF) {
        C = addScaled(C, C0Lin, -alpha);
        V F = mulv(K, C);
        lambdaLin = addScaled(F, lambdaLin, 1.0);
      }
      penaltyLin = addScaled(penaltyLin, vabs(C), solver->betaLin);
      penaltyLin = componentMin(penaltyLin, rmin(stiffnessLin, PENALTY_MAX));
    }
    if (lengthSq(penaltyAng) > 0) {
      M K = diag(penaltyAng.x, penaltyAng.y, penaltyAng.z);
      V C = evaluateAng();
      if (stiffnessAng == INF) {
        C = addScaled(C, C0Ang, -alpha);
        V F = mulv(K, C);
        lambdaAng = addScaled(F, lambdaAng, 1.0);
      }
      penaltyAng = addScaled(penaltyAng, vabs(C), solver->betaAng);
      penaltyAng = componentMin(penaltyAng, rmin(stiffnessAng, PENALTY_MAX));
    }
    if (lengthSq(lambdaAng) > fracture * fracture) {
      penaltyLin = penaltyAng = lambdaLin = lambdaAng = V();
      broken = true;
    }
  }
};

/** forces.ts:241-297. */
struct Spring : Force {
  V rA, rB;
  R rest, stiffness;
  Spring(V a, V b, R k, R r = -1) : rA(a), rB(b), rest(r), stiffness(k) {}
  int kind() const override { return 2; }
  void created() {
    if (rest < 0) {
      V pA = transform(bodyA->positionLin, bodyA->positionAng, rA);
      V pB = transform(bodyB->positionLin, bodyB->positionAng, rB);
      rest = length(sub(pA, pB));
    }
  }
  bool initialize() override { return true; }
  void updatePrimal(Rigid *body, R, BodySystem &sys) override {
    V pA = transform(bodyA->positionLin, bodyA->positionAng, rA);
    V pB = transform(bodyB->positionLin, bodyB->positionAng, rB);
    V d = sub(pA, pB);
    R dLen = length(d);
    if (dLen <= 1.0e-6) return;
    V n = divs(d, dLen);
    R c = dLen - rest;
    R f = stiffness * c;
    V jl, ja;
    if (body == bodyA) {
      V r = rotate(bodyA->positionAng, rA);
      jl = n;
      ja = cross(r, n);
    } else {
      V r = rotate(bodyB->positionAng, rB);
      jl = neg(n);
      ja = neg(cross(r, n));
    }
    addm(sys.lhsLin, scalem(outer(jl, jl), stiffness));
    addm(sys.lhsAng, scalem(outer(ja, ja), stiffness));
    addm(sys.lhsCross, scalem(outer(ja, jl), stiffness));
    sys.rhsLin = addScaled(sys.rhsLin, scale(jl, f), 1.0);
    sys.rhsAng = addScaled(sys.rhsAng, scale(ja, f), 1.0);
  }
  void updateDual(R) override {}
};

struct IgnoreCollision : Force {
  int kind() const override { return 3; }
  void created() {}
  bool initialize() override { return true; }
  void updatePrimal(Rigid *, R, BodySystem &) override {}
  void updateDual(R) override {}
};

/** solver.ts:134-219 (the package's CPU step when `package` is set). */
inline void Solver::step() {
  if (package) {
    packageStep();
    return;
  }
  // Broadphase: O(n²) bounding spheres, newest first
  for (int i = int(bodies.size()) - 1; i >= 0; i--) {
    Rigid *A = bodies[i];
    for (int j = i - 1; j >= 0; j--) {
      Rigid *B = bodies[j];
      V dp = sub(A->positionLin, B->positionLin);
      R r = A->radius + B->radius;
      if (dot(dp, dp) <= r * r && !A->constrainedTo(B)) addForce<Manifold>(A, B);
    }
  }
  // Initialize and warm start forces; drop the inactive ones
  std::vector<Force *> kept;
  for (Force *f : forces) {
    if (f->initialize()) kept.push_back(f);
    else f->unlinkFromBodies();
  }
  forces = kept;

  // Bodies
  for (int b = int(bodies.size()) - 1; b >= 0; b--) {
    Rigid *body = bodies[b];
    body->inertialLin = addScaled(body->positionLin, body->velocityLin, dt);
    if (body->mass > 0) body->inertialLin.z += gravity * (dt * dt);
    body->inertialAng = qadd(body->positionAng, scale(body->velocityAng, dt));
    V accel = divs(sub(body->velocityLin, body->prevVelocityLin), dt);
    R accelExt = accel.z * rsign(gravity);
    R accelWeight = rclamp(accelExt / std::fabs(gravity), 0.0, 1.0);
    if (!std::isfinite(accelWeight)) accelWeight = 0;
    body->initialLin = body->positionLin;
    body->initialAng = body->positionAng;
    if (body->mass > 0) {
      body->positionLin = addScaled(body->positionLin, body->velocityLin, dt);
      body->positionLin.z += gravity * (accelWeight * dt * dt);
      body->positionAng = qadd(body->positionAng, scale(body->velocityAng, dt));
    }
  }

  for (int it = 0; it < iterations; it++) {
    for (int b = int(bodies.size()) - 1; b >= 0; b--) {
      Rigid *body = bodies[b];
      if (body->mass <= 0) continue;
      BodySystem sys;
      R m = body->mass;
      V I = body->moment;
      sys.lhsLin = diag(m / (dt * dt), m / (dt * dt), m / (dt * dt));
      sys.lhsAng = diag(I.x / (dt * dt), I.y / (dt * dt), I.z / (dt * dt));
      sys.rhsLin = mulv(sys.lhsLin, sub(body->positionLin, body->inertialLin));
      sys.rhsAng = mulv(sys.lhsAng, qsub(body->positionAng, body->inertialAng));
      for (int f = int(body->forces.size()) - 1; f >= 0; f--) body->forces[f]->updatePrimal(body, alpha, sys);
      V dxLin, dxAng;
      solve6(sys.lhsLin, sys.lhsAng, sys.lhsCross, neg(sys.rhsLin), neg(sys.rhsAng), dxLin, dxAng);
      body->positionLin = addScaled(body->positionLin, dxLin, 1.0);
      body->positionAng = qadd(body->positionAng, dxAng);
    }
    for (Force *f : forces) f->updateDual(alpha);
  }

  for (Rigid *body : bodies) {
    body->prevVelocityLin = body->velocityLin;
    if (body->mass > 0) {
      body->velocityLin = divs(sub(body->positionLin, body->initialLin), dt);
      body->velocityAng = divs(qsub(body->positionAng, body->initialAng), dt);
    }
  }
}

/**
 * cpuBroadphase3D (cpuPhysics3D.js:304-363): a hash grid of cell = max(0.1, 2 × the median bounding radius of the first
 * 1,024 dynamic bodies); each body, in index order, is binned into every cell its bounding sphere's box touches and
 * paired with the bodies already there (a body touching more than 128 cells is paired with every body instead); a pair
 * (higher, lower index), not both static and not seen before, whose bounding spheres overlap and that no force links
 * yet, gets a new manifold.
 */
inline void Solver::packageBroadphase() {
  const size_t n = bodies.size();
  std::vector<R> radii;
  for (Rigid *b : bodies)
    if (b->mass > 0 && radii.size() < 1024) radii.push_back(b->radius);
  std::sort(radii.begin(), radii.end());
  const R cell = std::max(R(0.1), (radii.empty() ? R(1) : radii[radii.size() / 2]) * 2);
  // Cell lists in insertion order (the reference's Map of arrays keyed by "x,y,z") as open-addressing slots whose
  // entries chain head → tail; the seen-pair set as an open-addressing table of (hi · n + lo) + 1
  size_t slots = 1024;
  while (slots < 8 * n + 1024) slots *= 2;
  struct Cell {
    int64_t x, y, z;
    int32_t head = -1, tail = -1;
  };
  std::vector<Cell> cells(slots);
  std::vector<char> used(slots, 0);
  std::vector<int32_t> entryBody, entryNext;
  auto hashCell = [&](int64_t x, int64_t y, int64_t z) {
    uint64_t h = uint64_t(x) * 0x9E3779B97F4A7C15ull ^ uint64_t(y) * 0xC2B2AE3D27D4EB4Full ^ uint64_t(z) * 0x165667B19E3779F9ull;
    return size_t(h ^ (h >> 29)) & (slots - 1);
  };
  size_t seenSlots = 1024;
  while (seenSlots < 32 * n + 1024) seenSlots *= 2;
  std::vector<uint64_t> seen(seenSlots, 0);
  size_t seenCount = 0;
  auto firstTime = [&](uint64_t key) {
    if (2 * (seenCount + 1) > seenSlots) {  // grow (rehash) past half full
      std::vector<uint64_t> old;
      old.swap(seen);
      seenSlots *= 2;
      seen.assign(seenSlots, 0);
      for (uint64_t k : old)
        if (k) {
          size_t h = size_t((k * 0x9E3779B97F4A7C15ull) >> 20) & (seenSlots - 1);
          while (seen[h]) h = (h + 1) & (seenSlots - 1);
          seen[h] = k;
        }
    }
    const uint64_t k = key + 1;
    size_t h = size_t((k * 0x9E3779B97F4A7C15ull) >> 20) & (seenSlots - 1);
    while (seen[h]) {
      if (seen[h] == k) return false;
      h = (h + 1) & (seenSlots - 1);
    }
    seen[h] = k;
    seenCount++;
    return true;
  };
  std::vector<uint32_t> large;
  auto pair = [&](uint32_t a, uint32_t b) {
    if (a == b || (bodies[a]->mass <= 0 && bodies[b]->mass <= 0)) return;
    const uint32_t hi = std::max(a, b), lo = std::min(a, b);
    if (!firstTime(uint64_t(hi) * n + lo)) return;
    Rigid *A = bodies[hi], *B = bodies[lo];
    const R r = A->radius + B->radius;
    const V d = sub(A->positionLin, B->positionLin);
    if (dot(d, d) <= r * r && !A->constrainedTo(B)) addForce<Manifold>(A, B);
  };
  for (uint32_t i = 0; i < n; i++) {
    const Rigid *b = bodies[i];
    int64_t lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
      lo[k] = int64_t(std::floor((b->positionLin[k] - b->radius) / cell));
      hi[k] = int64_t(std::floor((b->positionLin[k] + b->radius) / cell));
    }
    if ((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) > 128) {
      large.push_back(i);
      continue;
    }
    for (int64_t x = lo[0]; x <= hi[0]; x++)
      for (int64_t y = lo[1]; y <= hi[1]; y++)
        for (int64_t z = lo[2]; z <= hi[2]; z++) {
          size_t h = hashCell(x, y, z);
          while (used[h] && !(cells[h].x == x && cells[h].y == y && cells[h].z == z)) h = (h + 1) & (slots - 1);
          Cell &c = cells[h];
          if (!used[h]) {
            used[h] = 1;
            c.x = x;
            c.y = y;
            c.z = z;
          }
          for (int32_t e = c.head; e >= 0; e = entryNext[size_t(e)]) pair(i, uint32_t(entryBody[size_t(e)]));
          const int32_t e = int32_t(entryBody.size());
          entryBody.push_back(int32_t(i));
          entryNext.push_back(-1);
          if (c.tail >= 0) entryNext[size_t(c.tail)] = e;
          else c.head = e;
          c.tail = e;
        }
  }
  for (uint32_t i : large)
    for (uint32_t j = 0; j < n; j++) pair(i, j);
}

/**
 * CpuSolver3D.step (cpuSolver3D.ts:144-280) for boxes, joints and springs: broadphase, force initialisation (dropping
 * the inactive), body warm start with gravity along `up` and the adaptive weight, `iterations` × (each dynamic body's
 * Newton step, newest first, its forces in reverse order; every force's dual), BDF1 velocities. The colour-parallel
 * variant takes the Newton steps colour by colour (greedy colouring of the bodies linked by forces, in index order),
 * a colour's bodies in parallel.
 */
inline void Solver::packageStep() {
  packageBroadphase();
  std::vector<char> alive(forces.size());
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads) if (threads > 1)
  for (size_t k = 0; k < forces.size(); k++) alive[k] = forces[k]->initialize() ? 1 : 0;
  std::vector<Force *> kept;
  for (size_t k = 0; k < forces.size(); k++) {
    if (alive[k]) kept.push_back(forces[k]);
    else forces[k]->unlinkFromBodies();
  }
  forces = kept;

  const R g = gravity;
  for (int bi = int(bodies.size()) - 1; bi >= 0; bi--) {
    Rigid *body = bodies[size_t(bi)];
    body->inertialLin = addScaled(body->positionLin, body->velocityLin, dt);
    if (body->mass > 0) body->inertialLin = addScaled(body->inertialLin, up, g * dt * dt);
    body->inertialAng = qadd(body->positionAng, scale(body->velocityAng, dt));
    const V accel = divs(sub(body->velocityLin, body->prevVelocityLin), dt);
    const R accelExt = dot(accel, up) * rsign(g);
    R accelWeight = rclamp(accelExt / std::fabs(g), 0.0, 1.0);
    if (!std::isfinite(accelWeight)) accelWeight = 0;
    body->initialLin = body->positionLin;
    body->initialAng = body->positionAng;
    if (body->mass > 0) {
      body->positionLin = addScaled(body->positionLin, body->velocityLin, dt);
      body->positionLin = addScaled(body->positionLin, up, g * accelWeight * dt * dt);
      body->positionAng = qadd(body->positionAng, scale(body->velocityAng, dt));
    }
  }

  auto newton = [&](Rigid *body) {
    BodySystem sys;
    const R m = body->mass;
    const V I = body->moment;
    sys.lhsLin = diag(m / (dt * dt), m / (dt * dt), m / (dt * dt));
    sys.lhsAng = diag(I.x / (dt * dt), I.y / (dt * dt), I.z / (dt * dt));
    sys.rhsLin = mulv(sys.lhsLin, sub(body->positionLin, body->inertialLin));
    sys.rhsAng = mulv(sys.lhsAng, qsub(body->positionAng, body->inertialAng));
    for (int f = int(body->forces.size()) - 1; f >= 0; f--) body->forces[size_t(f)]->updatePrimal(body, alpha, sys);
    V dxLin, dxAng;
    solve6(sys.lhsLin, sys.lhsAng, sys.lhsCross, neg(sys.rhsLin), neg(sys.rhsAng), dxLin, dxAng);
    body->positionLin = addScaled(body->positionLin, dxLin, 1.0);
    body->positionAng = qadd(body->positionAng, dxAng);
  };
  std::vector<std::vector<Rigid *>> byColor;
  if (threads > 1) {  // greedy colours in index order over the bodies the live forces link
    std::unordered_map<const Rigid *, uint32_t> index;
    for (uint32_t i = 0; i < bodies.size(); i++) index[bodies[i]] = i;
    colors.assign(bodies.size(), ~0u);
    for (uint32_t i = 0; i < bodies.size(); i++) {
      Rigid *b = bodies[i];
      if (b->mass <= 0) continue;
      std::vector<char> used;
      for (const Force *f : b->forces) {
        const Rigid *o = f->bodyA == b ? f->bodyB : f->bodyA;
        if (!o || o->mass <= 0) continue;
        const uint32_t c = colors[index[o]];
        if (c == ~0u) continue;
        if (used.size() <= c) used.resize(c + 1, 0);
        used[c] = 1;
      }
      uint32_t c = 0;
      while (c < used.size() && used[c]) c++;
      colors[i] = c;
      if (byColor.size() <= c) byColor.resize(c + 1);
      byColor[c].push_back(b);
    }
  }
  for (int it = 0; it < iterations; it++) {
    if (threads > 1) {
      for (auto &group : byColor) {
#pragma omp parallel for schedule(dynamic, 32) num_threads(threads)
        for (size_t k = 0; k < group.size(); k++) newton(group[k]);
      }
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
      for (size_t k = 0; k < forces.size(); k++) forces[k]->updateDual(alpha);
    } else {
      for (int bi = int(bodies.size()) - 1; bi >= 0; bi--)
        if (bodies[size_t(bi)]->mass > 0) newton(bodies[size_t(bi)]);
      for (Force *f : forces) f->updateDual(alpha);
    }
  }
  for (Rigid *body : bodies) {
    body->prevVelocityLin = body->velocityLin;
    if (body->mass > 0) {
      body->velocityLin = divs(sub(body->positionLin, body->initialLin), dt);
      body->velocityAng = divs(qsub(body->positionAng, body->initialAng), dt);
    }
  }
}

/** Builds scene `index` of scenes.ts:236-251 into an empty solver; returns false for an unknown index. */
bool buildScene(Solver &s, int index);
extern const char *const SCENE_NAMES[14];

}  // namespace ref
}  // namespace oracle

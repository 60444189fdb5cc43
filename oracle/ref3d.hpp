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
    Found<R> found = collideBoxes(bA, bB, false, 16, sat);
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

/** solver.ts:134-219. */
inline void Solver::step() {
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

/** Builds scene `index` of scenes.ts:236-251 into an empty solver; returns false for an unknown index. */
bool buildScene(Solver &s, int index);
extern const char *const SCENE_NAMES[14];

}  // namespace ref
}  // namespace oracle

// ORACLE — test infrastructure only (see math.hpp).
//
// The 14 upstream demo scenes (reference/three-avbd/src/avbd3d/ref/scenes.ts:1-258, z up) and a C interface to the
// FP64 upstream restatement for the tests (ctypes).
#include <cstring>

#include "ref3d.hpp"

namespace oracle {
namespace ref {

const char *const SCENE_NAMES[14] = {"Empty",       "Ground", "Dynamic Friction", "Static Friction", "Pyramid",
                                     "Rope",        "Heavy Rope", "Spring",       "Spring Ratio",    "Stack",
                                     "Stack Ratio", "Soft Body",  "Bridge",       "Breakable"};

static Rigid *ground(Solver &s, R z = 0) { return s.addBody(V(100, 100, 1), 0, 0.5, V(0, 0, z)); }

bool buildScene(Solver &s, int index) {
  switch (index) {
    case 0:  // scenes.ts:11
      return true;
    case 1:  // scenes.ts:13-17
      ground(s);
      s.addBody(V(1, 1, 1), 1, 0.5, V(0, 0, 4));
      return true;
    case 2:  // scenes.ts:19-25
      ground(s);
      for (int x = 0; x <= 10; x++) s.addBody(V(1, 1, 0.5), 1, 5 - (x / 10.0) * 5, V(0, -30 + x * 2, 0.75), V(10, 0, 0));
      return true;
    case 3: {  // scenes.ts:27-46
      ground(s);
      R angle = 30 * 0.017453292519943295;
      Rigid *ramp = s.addBody(V(40, 24, 1), 0, 1, V(0, 0, 3));
      ramp->positionAng = Q(0, std::sin(angle * 0.5), 0, std::cos(angle * 0.5));
      V tangent = normalize(rotate(ramp->positionAng, V(1, 0, 0)));
      V normal = normalize(rotate(ramp->positionAng, V(0, 0, 1)));
      for (int i = 0; i <= 10; i++) {
        R friction = (i / 10.0) * 0.25 + 0.25;
        R y = -10 + i * 2;
        V pos = addScaled(ramp->positionLin, tangent, -12.0);
        pos.y += y;
        pos = addScaled(pos, normal, 1.05);
        s.addBody(V(1, 1, 1), 1, friction, pos);
      }
      return true;
    }
    case 4: {  // scenes.ts:49-57
      const int size = 16;
      ground(s, -0.5);
      for (int y = 0; y < size; y++)
        for (int x = 0; x < size - y; x++)
          s.addBody(V(1, 0.5, 0.5), 1, 0.5, V(x * 1.01 + y * 0.5 - size / 2.0, 0, y * 0.85 + 0.5));
      return true;
    }
    case 5: {  // scenes.ts:59-68
      ground(s, -20);
      Rigid *prev = nullptr;
      for (int i = 0; i < 20; i++) {
        Rigid *curr = s.addBody(V(1, 0.5, 0.5), i == 0 ? 0 : 1, 0.5, V(i, 0, 10));
        if (prev) s.addForce<Joint>(prev, curr, V(0.5, 0, 0), V(-0.5, 0, 0));
        prev = curr;
      }
      return true;
    }
    case 6: {  // scenes.ts:70-83
      const int N = 20;
      const R SIZE = 5;
      ground(s, -20);
      Rigid *prev = nullptr;
      for (int i = 0; i < N; i++) {
        bool last = i == N - 1;
        Rigid *curr = s.addBody(last ? V(SIZE, SIZE, SIZE) : V(1, 0.5, 0.5), i == 0 ? 0 : 1, 0.5,
                                V(i + (last ? SIZE / 2 : 0), 0, 10));
        if (prev) s.addForce<Joint>(prev, curr, V(0.5, 0, 0), last ? V(-SIZE / 2, 0, 0) : V(-0.5, 0, 0));
        prev = curr;
      }
      return true;
    }
    case 7: {  // scenes.ts:85-91
      ground(s);
      Rigid *anchor = s.addBody(V(1, 1, 1), 0, 0.5, V(0, 0, 14));
      Rigid *block = s.addBody(V(2, 2, 2), 1, 0.5, V(0, 0, 8));
      s.addForce<Spring>(anchor, block, V(0, 0, 0), V(0, 0, 0), 100.0, 4.0);
      return true;
    }
    case 8: {  // scenes.ts:93-104
      const int N = 8;
      ground(s, -10);
      Rigid *prev = nullptr;
      for (int i = 0; i < N; i++) {
        R x = (i - (N - 1) * 0.5) * 3;
        Rigid *curr = s.addBody(V(1, 0.75, 0.75), (i == 0 || i == N - 1) ? 0 : 1, 0.5, V(x, 0, 12));
        if (prev) s.addForce<Spring>(prev, curr, V(0.5, 0, 0), V(-0.5, 0, 0), i % 2 == 0 ? 10.0 : 10000.0, 3.0);
        prev = curr;
      }
      return true;
    }
    case 9:  // scenes.ts:106-110
      ground(s);
      for (int i = 0; i < 10; i++) s.addBody(V(1, 1, 1), 1, 0.5, V(0, 0, i * 1.5 + 1));
      return true;
    case 10: {  // scenes.ts:112-125
      const R groundThickness = 1;
      s.addBody(V(100, 100, groundThickness), 0, 0.5, V(0, 0, 0));
      R topZ = groundThickness * 0.5;
      R sz = 1;
      for (int i = 0; i < 4; i++) {
        R half = sz * 0.5;
        R centerZ = topZ + half;
        s.addBody(V(sz, sz, sz), 1, 0.5, V(0, 0, centerZ));
        topZ = centerZ + half;
        sz *= 2;
      }
      return true;
    }
    case 11: {  // scenes.ts:127-184
      ground(s);
      const R Klin = 1000, Kang = 250;
      const int W = 4, D = 4, H = 4, N = 3;
      const R size = 0.8, half = size * 0.5, baseZ = 8, stackGap = 2;
      for (int i = 0; i < N; i++) {
        R stackZ = i * (H * size + stackGap);
        Rigid *grid[4][4][4];
        for (int x = 0; x < W; x++)
          for (int y = 0; y < D; y++)
            for (int z = 0; z < H; z++) {
              R px = (x - (W - 1) * 0.5) * size;
              R py = (y - (D - 1) * 0.5) * size;
              R pz = baseZ + stackZ + z * size;
              grid[x][y][z] = s.addBody(V(size, size, size), 1, 0.5, V(px, py, pz));
            }
        for (int x = 1; x < W; x++)
          for (int y = 0; y < D; y++)
            for (int z = 0; z < H; z++)
              s.addForce<Joint>(grid[x - 1][y][z], grid[x][y][z], V(half, 0, 0), V(-half, 0, 0), Klin, Kang);
        for (int x = 0; x < W; x++)
          for (int y = 1; y < D; y++)
            for (int z = 0; z < H; z++)
              s.addForce<Joint>(grid[x][y - 1][z], grid[x][y][z], V(0, half, 0), V(0, -half, 0), Klin, Kang);
        for (int x = 0; x < W; x++)
          for (int y = 0; y < D; y++)
            for (int z = 1; z < H; z++)
              s.addForce<Joint>(grid[x][y][z - 1], grid[x][y][z], V(0, 0, half), V(0, 0, -half), Klin, Kang);
        for (int x = 1; x < W; x++)
          for (int y = 0; y < D; y++)
            for (int z = 1; z < H; z++) {
              s.addForce<IgnoreCollision>(grid[x - 1][y][z - 1], grid[x][y][z]);
              s.addForce<IgnoreCollision>(grid[x][y][z - 1], grid[x - 1][y][z]);
            }
        for (int x = 0; x < W; x++)
          for (int y = 1; y < D; y++)
            for (int z = 1; z < H; z++) {
              s.addForce<IgnoreCollision>(grid[x][y - 1][z - 1], grid[x][y][z]);
              s.addForce<IgnoreCollision>(grid[x][y][z - 1], grid[x][y - 1][z]);
            }
        for (int x = 1; x < W; x++)
          for (int y = 1; y < D; y++)
            for (int z = 0; z < H; z++) {
              s.addForce<IgnoreCollision>(grid[x - 1][y - 1][z], grid[x][y][z]);
              s.addForce<IgnoreCollision>(grid[x][y - 1][z], grid[x - 1][y][z]);
            }
      }
      return true;
    }
    case 12: {  // scenes.ts:186-212
      const int N = 40;
      const R plankLength = 1, plankWidth = 4, plankHeight = 0.5;
      const R halfLength = plankLength * 0.5, halfWidth = plankWidth * 0.5;
      ground(s);
      Rigid *prev = nullptr;
      for (int i = 0; i < N; i++) {
        Rigid *curr = s.addBody(V(plankLength, plankWidth, plankHeight), (i == 0 || i == N - 1) ? 0 : 1, 0.5,
                                V(i - N / 2.0, 0, 10));
        if (prev) {
          s.addForce<Joint>(prev, curr, V(halfLength, halfWidth, 0), V(-halfLength, halfWidth, 0), INF, 0.0);
          s.addForce<Joint>(prev, curr, V(halfLength, -halfWidth, 0), V(-halfLength, -halfWidth, 0), INF, 0.0);
        }
        prev = curr;
      }
      for (int x = 0; x < N / 4; x++)
        for (int y = 0; y < N / 8; y++) s.addBody(V(1, 1, 1), 1, 0.5, V(x - N / 8.0, 0, y + 12));
      return true;
    }
    case 13: {  // scenes.ts:214-232
      const int N = 10, M = 5;
      const R breakForce = 90;
      ground(s);
      Rigid *prev = nullptr;
      for (int i = 0; i <= N; i++) {
        Rigid *curr = s.addBody(V(1, 1, 0.5), 1, 0.5, V(i - N / 2.0, 0, 6));
        if (prev) s.addForce<Joint>(prev, curr, V(0.5, 0, 0), V(-0.5, 0, 0), INF, INF, breakForce);
        prev = curr;
      }
      s.addBody(V(1, 1, 5), 0, 0.5, V(-N / 2.0, 0, 2.5));
      s.addBody(V(1, 1, 5), 0, 0.5, V(N / 2.0, 0, 2.5));
      for (int i = 0; i < M; i++) s.addBody(V(2, 1, 1), 1, 0.5, V(0, 0, i * 2 + 8));
      return true;
    }
  }
  return false;
}

}  // namespace ref
}  // namespace oracle

using namespace oracle::ref;

extern "C" {

void *ref_create(int scene, int iterations) {
  auto *s = new Solver();
  if (iterations > 0) s->iterations = iterations;
  if (!buildScene(*s, scene)) {
    delete s;
    return nullptr;
  }
  return s;
}
void ref_destroy(void *h) { delete static_cast<Solver *>(h); }
void ref_step(void *h, int n) {
  auto *s = static_cast<Solver *>(h);
  for (int i = 0; i < n; i++) s->step();
}
int ref_body_count(void *h) { return int(static_cast<Solver *>(h)->bodies.size()); }
int ref_force_count(void *h) { return int(static_cast<Solver *>(h)->forces.size()); }
/** Per body: pos xyz, quat xyzw, vel xyz, angVel xyz, mass, moment xyz, radius (17 doubles). */
void ref_read_bodies(void *h, double *out) {
  auto *s = static_cast<Solver *>(h);
  for (size_t i = 0; i < s->bodies.size(); i++) {
    Rigid *b = s->bodies[i];
    double *o = out + 17 * i;
    o[0] = b->positionLin.x; o[1] = b->positionLin.y; o[2] = b->positionLin.z;
    o[3] = b->positionAng.x; o[4] = b->positionAng.y; o[5] = b->positionAng.z; o[6] = b->positionAng.w;
    o[7] = b->velocityLin.x; o[8] = b->velocityLin.y; o[9] = b->velocityLin.z;
    o[10] = b->velocityAng.x; o[11] = b->velocityAng.y; o[12] = b->velocityAng.z;
    o[13] = b->mass; o[14] = b->moment.x; o[15] = b->moment.y; o[16] = b->moment.z;
  }
}
/** Number of live joints (Joint forces) and the largest |linear joint error| among them. */
int ref_joint_stats(void *h, double *maxLinError) {
  auto *s = static_cast<Solver *>(h);
  int n = 0;
  double err = 0;
  for (Force *f : s->forces)
    if (f->kind() == 1) {
      n++;
      double e = length(static_cast<Joint *>(f)->evaluateLin());
      if (e > err) err = e;
    }
  if (maxLinError) *maxLinError = err;
  return n;
}
/** Ray pick of the dynamic bodies (solver.ts:89-122); returns the body index or -1. */
int ref_pick(void *h, const double *origin, const double *dir, double *local) {
  auto *s = static_cast<Solver *>(h);
  const double eps = 1.0e-6;
  int best = -1;
  double bestT = 0;
  V bestLocal;
  V org(origin[0], origin[1], origin[2]), dv(dir[0], dir[1], dir[2]);
  for (int bi = int(s->bodies.size()) - 1; bi >= 0; bi--) {
    Rigid *body = s->bodies[bi];
    if (body->mass <= 0) continue;
    V o = rotateInv(body->positionAng, sub(org, body->positionLin));
    V d = rotateInv(body->positionAng, dv);
    double tEnter = 0, tExit = INF;
    bool hit = true;
    for (int i = 0; i < 3 && hit; i++) {
      double halfw = body->size[i] * 0.5;
      if (std::fabs(d[i]) < eps) {
        if (o[i] < -halfw || o[i] > halfw) hit = false;
        continue;
      }
      double invD = 1 / d[i];
      double t0 = (-halfw - o[i]) * invD, t1 = (halfw - o[i]) * invD;
      if (t0 > t1) std::swap(t0, t1);
      tEnter = std::max(tEnter, t0);
      tExit = std::min(tExit, t1);
      if (tEnter > tExit) hit = false;
    }
    if (!hit) continue;
    double tHit = tEnter >= 0 ? tEnter : tExit;
    if (tHit < 0) continue;
    if (best < 0 || tHit < bestT) {
      best = bi;
      bestT = tHit;
      bestLocal = addScaled(o, d, tHit);
    }
  }
  if (best >= 0) {
    local[0] = bestLocal.x;
    local[1] = bestLocal.y;
    local[2] = bestLocal.z;
  }
  return best;
}
/** Adds a world-anchored joint (bodyA = world) to body b; returns nothing. Stiffness < 0 means infinite. */
void ref_add_world_joint(void *h, int b, const double *target, const double *local, double kLin, double kAng) {
  auto *s = static_cast<Solver *>(h);
  s->addForce<Joint>(nullptr, s->bodies[b], V(target[0], target[1], target[2]), V(local[0], local[1], local[2]),
                     kLin < 0 ? INF : kLin, kAng < 0 ? INF : kAng);
}
/** |linear error| of a fresh world joint between target and body b's local point (forces.ts evaluateLin). */
double ref_world_joint_error(void *h, int b, const double *target, const double *local) {
  auto *s = static_cast<Solver *>(h);
  Rigid *body = s->bodies[b];
  V pB = transform(body->positionLin, body->positionAng, V(local[0], local[1], local[2]));
  return length(sub(V(target[0], target[1], target[2]), pB));
}
/**
 * The package's CPU path (CpuSolver3D.step) over packed scene records: bodies 40 floats each (boxes only: a sphere or
 * capsule makes it return NULL), joints 32 floats + info (type 1 joint / 2 spring, a or -1, b); params 16 floats
 * (dt, gravity, betaLin, betaAng, gamma, alpha, up xyz, flags bits: 2 face bias, 16 mass penalty). threads > 1: the
 * colour-parallel variant.
 */
void *cpu_create(uint32_t nBodies, const float *bodyRecs, uint32_t nJoints, const float *jointRecs,
                 const int32_t *info, const float *params, int iterations, int threads) {
  auto s = std::make_unique<Solver>();
  s->package = true;
  s->dt = params[0];
  s->gravity = params[1];
  s->betaLin = params[2];
  s->betaAng = params[3];
  s->gamma = params[4];
  s->alpha = params[5];
  s->up = V(params[6], params[7], params[8]);
  uint32_t flags;
  std::memcpy(&flags, params + 9, 4);
  s->faceBias = (flags & 2u) != 0;
  s->massPenalty = (flags & 16u) != 0;
  s->iterations = iterations;
  s->threads = threads < 1 ? 1 : threads;
  for (uint32_t i = 0; i < nBodies; i++) {
    const float *r = bodyRecs + 40ull * i;
    if (r[39] != 0.f) return nullptr;  // boxes only
    Rigid *b = s->addBody(V(r[16], r[17], r[18]), 0.0, r[3], V(r[0], r[1], r[2]), V(r[32], r[33], r[34]));
    b->positionAng = Q(r[4], r[5], r[6], r[7]);
    b->mass = r[19];
    b->moment = V(r[20], r[21], r[22]);
    b->radius = r[23];
    b->velocityAng = V(r[36], r[37], r[38]);
  }
  auto fin = [](float x) { return x >= 1e30f ? INF : double(x); };
  for (uint32_t j = 0; j < nJoints; j++) {
    const float *k = jointRecs + 32ull * j;
    const int32_t *f = info + 4ull * j;
    Rigid *a = f[1] >= 0 ? s->bodies[size_t(f[1])] : nullptr;
    Rigid *b = s->bodies[size_t(f[2])];
    if (f[0] == 1)
      s->addForce<Joint>(a, b, V(k[24], k[25], k[26]), V(k[28], k[29], k[30]), fin(k[3]), fin(k[7]), fin(k[11]));
    else if (f[0] == 2)
      s->addForce<Spring>(a, b, V(k[24], k[25], k[26]), V(k[28], k[29], k[30]), double(k[3]), double(k[27]));
    else
      return nullptr;
  }
  return s.release();
}
void cpu_destroy(void *h) { delete static_cast<Solver *>(h); }
void cpu_step(void *h, int n) {
  auto *s = static_cast<Solver *>(h);
  for (int i = 0; i < n; i++) s->step();
}
void cpu_set_threads(void *h, int threads) { static_cast<Solver *>(h)->threads = threads < 1 ? 1 : threads; }
/** Bodies as 40 values each in the record layout (pos, friction, rot, size, mass, moment, radius, vel, angVel). */
void cpu_read_bodies(void *h, double *out) {
  auto *s = static_cast<Solver *>(h);
  for (size_t i = 0; i < s->bodies.size(); i++) {
    const Rigid *b = s->bodies[i];
    double *o = out + 40 * i;
    for (int k = 0; k < 40; k++) o[k] = 0;
    o[0] = b->positionLin.x; o[1] = b->positionLin.y; o[2] = b->positionLin.z; o[3] = b->friction;
    o[4] = b->positionAng.x; o[5] = b->positionAng.y; o[6] = b->positionAng.z; o[7] = b->positionAng.w;
    o[16] = b->size.x; o[17] = b->size.y; o[18] = b->size.z; o[19] = b->mass;
    o[20] = b->moment.x; o[21] = b->moment.y; o[22] = b->moment.z; o[23] = b->radius;
    o[32] = b->velocityLin.x; o[33] = b->velocityLin.y; o[34] = b->velocityLin.z;
    o[36] = b->velocityAng.x; o[37] = b->velocityAng.y; o[38] = b->velocityAng.z;
  }
}
/** Live forces by kind: out[0] manifolds, out[1] joints, out[2] springs; returns the contact points. */
uint32_t cpu_counts(void *h, uint32_t *out) {
  auto *s = static_cast<Solver *>(h);
  uint32_t points = 0;
  out[0] = out[1] = out[2] = 0;
  for (const Force *f : s->forces) {
    const int k = f->kind();
    if (k < 3) out[k]++;
    if (k == 0) points += uint32_t(static_cast<const Manifold *>(f)->contacts.size());
  }
  return points;
}
/** The broadphase alone (manifolds created for the new pairs, not initialised), for checks of its pair set. */
void cpu_broadphase(void *h) { static_cast<Solver *>(h)->packageBroadphase(); }
/** The manifold pairs (a > b as body indices), for checks of the broadphase. */
uint32_t cpu_read_pairs(void *h, uint32_t *out) {
  auto *s = static_cast<Solver *>(h);
  std::unordered_map<const Rigid *, uint32_t> index;
  for (uint32_t i = 0; i < s->bodies.size(); i++) index[s->bodies[i]] = i;
  uint32_t n = 0;
  for (const Force *f : s->forces)
    if (f->kind() == 0) {
      if (out) {
        const uint32_t a = index[f->bodyA], b = index[f->bodyB];
        out[2 * n] = std::max(a, b);
        out[2 * n + 1] = std::min(a, b);
      }
      n++;
    }
  return n;
}

/** Mass properties of a box of the given size and density (body.ts:38-47): mass, moment xyz, radius. */
void ref_box_mass(const double *size, double density, double *out) {
  Solver s;
  Rigid *b = s.addBody(V(size[0], size[1], size[2]), density, 0.5, V());
  out[0] = b->mass;
  out[1] = b->moment.x;
  out[2] = b->moment.y;
  out[3] = b->moment.z;
  out[4] = b->radius;
}
}

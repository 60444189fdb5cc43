// ORACLE — test infrastructure only (see math.hpp).
//
// C interface (for the tests, through ctypes) to the GPU-semantics oracle: build a world from the reference's packed
// records, seed its persistent state from a state read back from the HIP library, step it in FP32 (the mirror) or
// FP64 (with the FP32 mirror's contacts and colours, to state the FP32 error of the solve), and read everything back in
// the same record layouts.
#include "queries.hpp"
#include <cstring>
#include <memory>

#include <omp.h>

#include "gpusem.hpp"

namespace {
struct OneThreadByDefault {
  OneThreadByDefault() { omp_set_num_threads(1); }
} oneThreadByDefault;
}  // namespace

using namespace oracle;
using namespace oracle::gpu;

namespace {

struct Handle {
  int precision;  // 32 or 64
  std::unique_ptr<World<float>> f;
  std::unique_ptr<World<double>> d;
};

template <class R> void unpackBody(const float *s, Body<R> &b) {
  b.pos = V3<R>(s[0], s[1], s[2]);
  b.friction = s[3];
  b.rot = Q4<R>(s[4], s[5], s[6], s[7]);
  b.ipos = V3<R>(s[8], s[9], s[10]);
  b.iposw = s[11];
  b.irot = Q4<R>(s[12], s[13], s[14], s[15]);
  b.size = V3<R>(s[16], s[17], s[18]);
  b.mass = s[19];
  b.moment = V3<R>(s[20], s[21], s[22]);
  b.radius = s[23];
  b.inertialPos = V3<R>(s[24], s[25], s[26]);
  b.stamp = floatToBits(s[27]);
  b.inertialRot = Q4<R>(s[28], s[29], s[30], s[31]);
  b.vel = V3<R>(s[32], s[33], s[34]);
  b.prevUp = s[35];
  b.angVel = V3<R>(s[36], s[37], s[38]);
  b.shape = s[39];
}
template <class R, class T> void packBody(const Body<R> &b, T *o) {
  const R v[40] = {b.pos.x, b.pos.y, b.pos.z, b.friction, b.rot.x, b.rot.y, b.rot.z, b.rot.w,
                   b.ipos.x, b.ipos.y, b.ipos.z, b.iposw, b.irot.x, b.irot.y, b.irot.z, b.irot.w,
                   b.size.x, b.size.y, b.size.z, b.mass, b.moment.x, b.moment.y, b.moment.z, b.radius,
                   b.inertialPos.x, b.inertialPos.y, b.inertialPos.z, 0, b.inertialRot.x, b.inertialRot.y,
                   b.inertialRot.z, b.inertialRot.w, b.vel.x, b.vel.y, b.vel.z, b.prevUp,
                   b.angVel.x, b.angVel.y, b.angVel.z, b.shape};
  for (int k = 0; k < 40; k++) o[k] = T(v[k]);
  o[27] = sizeof(T) == 4 ? T(bitsToFloat(b.stamp)) : T(b.stamp);
}
/** Word `word` (0..31) of a constraint record. */
template <class R> R *jointWord(Joint<R> &k, uint32_t word) {
  V3<R> *v[8] = {&k.penLin, &k.penAng, &k.lamLin, &k.lamAng, &k.c0Lin, &k.c0Ang, &k.rA, &k.rB};
  R *w[8] = {&k.stiffLin, &k.stiffAng, &k.frac, &k.torqueArm, &k.c0LinW, &k.c0AngW, &k.rest, &k.rBW};
  const uint32_t q = word / 4, c = word % 4;
  if (c == 3) return w[q];
  return c == 0 ? &v[q]->x : (c == 1 ? &v[q]->y : &v[q]->z);
}

template <class R> void unpackJoint(const float *s, const int32_t *info, Joint<R> &k) {
  k.penLin = V3<R>(s[0], s[1], s[2]); k.stiffLin = s[3];
  k.penAng = V3<R>(s[4], s[5], s[6]); k.stiffAng = s[7];
  k.lamLin = V3<R>(s[8], s[9], s[10]); k.frac = s[11];
  k.lamAng = V3<R>(s[12], s[13], s[14]); k.torqueArm = s[15];
  k.c0Lin = V3<R>(s[16], s[17], s[18]); k.c0LinW = s[19];
  k.c0Ang = V3<R>(s[20], s[21], s[22]); k.c0AngW = s[23];
  k.rA = V3<R>(s[24], s[25], s[26]); k.rest = s[27];
  k.rB = V3<R>(s[28], s[29], s[30]); k.rBW = s[31];
  k.type = info[0]; k.a = info[1]; k.b = info[2];
}
template <class R, class T> void packJoint(const Joint<R> &k, T *o) {
  const R v[32] = {k.penLin.x, k.penLin.y, k.penLin.z, k.stiffLin, k.penAng.x, k.penAng.y, k.penAng.z, k.stiffAng,
                   k.lamLin.x, k.lamLin.y, k.lamLin.z, k.frac, k.lamAng.x, k.lamAng.y, k.lamAng.z, k.torqueArm,
                   k.c0Lin.x, k.c0Lin.y, k.c0Lin.z, k.c0LinW, k.c0Ang.x, k.c0Ang.y, k.c0Ang.z, k.c0AngW,
                   k.rA.x, k.rA.y, k.rA.z, k.rest, k.rB.x, k.rB.y, k.rB.z, k.rBW};
  for (int i = 0; i < 32; i++) o[i] = T(v[i]);
}
template <class R> void unpackManifold(const uint32_t *w, Manifold<R> &m) {
  m.a = w[0]; m.b = w[1]; m.first = w[2]; m.countw = w[3];
  m.normal = V3<R>(bitsToFloat(w[4]), bitsToFloat(w[5]), bitsToFloat(w[6]));
  m.friction = bitsToFloat(w[7]);
}
template <class R> void packManifold(const Manifold<R> &m, uint32_t *w) {
  w[0] = m.a; w[1] = m.b; w[2] = m.first; w[3] = m.countw;
  w[4] = floatToBits(float(m.normal.x)); w[5] = floatToBits(float(m.normal.y)); w[6] = floatToBits(float(m.normal.z));
  w[7] = floatToBits(float(m.friction));
}
template <class R> void unpackContact(const uint32_t *w, Contact<R> &c) {
  auto f = [&](int i) { return R(bitsToFloat(w[i])); };
  c.rA = V3<R>(f(0), f(1), f(2)); c.key = w[3];
  c.rB = V3<R>(f(4), f(5), f(6)); c.c0x = f(7);
  c.pen = V3<R>(f(8), f(9), f(10)); c.c0y = f(11);
  c.lam = V3<R>(f(12), f(13), f(14)); c.c0z = f(15);
}
/** 16 values: rA, key (as its integer value), rB, c0x, pen, c0y, lam, c0z. */
template <class R, class T> void packContact(const Contact<R> &c, T *o) {
  const R v[16] = {c.rA.x, c.rA.y, c.rA.z, 0, c.rB.x, c.rB.y, c.rB.z, c.c0x,
                   c.pen.x, c.pen.y, c.pen.z, c.c0y, c.lam.x, c.lam.y, c.lam.z, c.c0z};
  for (int i = 0; i < 16; i++) o[i] = T(v[i]);
}

template <class R>
void build(World<R> &w, uint32_t nBodies, const float *bodyRecs, const uint32_t *bodyWorld, uint32_t nWorlds,
           const float *worldParams, uint32_t nJoints, const float *jointRecs, const int32_t *info, uint32_t nNoCollide,
           const uint32_t *noCollide, const uint32_t *filters) {
  w.bodies.resize(nBodies);
  w.restitution.assign(nBodies, R(0));
  w.world.assign(nBodies, 0);
  for (uint32_t i = 0; i < nBodies; i++) {
    unpackBody(bodyRecs + 40 * i, w.bodies[i]);
    if (bodyWorld) w.world[i] = bodyWorld[i];
  }
  w.params.resize(nWorlds);
  w.worldBase.assign(nWorlds, 0);
  for (uint32_t k = 0; k < nWorlds; k++) {
    const float *p = worldParams + 16 * k;
    WorldParams &q = w.params[k];
    q.dt = p[0]; q.gravity = p[1]; q.betaLin = p[2]; q.betaAng = p[3]; q.gamma = p[4]; q.alpha = p[5];
    q.up[0] = p[6]; q.up[1] = p[7]; q.up[2] = p[8];
    q.flags = floatToBits(p[9]);
    q.step = floatToBits(p[10]);
  }
  for (uint32_t i = nBodies; i-- > 0;) w.worldBase[w.world[i]] = i;
  w.joints.resize(nJoints);
  for (uint32_t j = 0; j < nJoints; j++) unpackJoint(jointRecs + 32 * j, info + 4 * j, w.joints[j]);
  w.noCollide.clear();
  for (uint32_t e = 0; e < nNoCollide; e++)
    w.noCollide.emplace_back(noCollide[3 * e], noCollide[3 * e + 1], noCollide[3 * e + 2]);
  std::sort(w.noCollide.begin(), w.noCollide.end());
  w.filters.assign(2 * size_t(nBodies), 0xffffffffu);
  if (filters) std::memcpy(w.filters.data(), filters, 8 * size_t(nBodies));
  w.refs.assign(nBodies, RefPose());
  w.colorA.assign(nBodies, NO_COLOR);
}

template <class R>
void setState(World<R> &w, const float *refs, const uint32_t *colors, uint32_t nManifolds, const uint32_t *manifolds,
              uint32_t nContacts, const uint32_t *contacts, const uint32_t *steps) {
  if (refs)
    for (size_t i = 0; i < w.n(); i++) {
      std::memcpy(w.refs[i].pos, refs + 8 * i, 16);
      std::memcpy(w.refs[i].rot, refs + 8 * i + 4, 16);
    }
  if (colors) std::memcpy(w.colorA.data(), colors, 4 * w.n());
  w.manifolds.resize(nManifolds);
  for (uint32_t m = 0; m < nManifolds; m++) unpackManifold(manifolds + 8 * m, w.manifolds[m]);
  w.contacts.resize(nContacts);
  for (uint32_t c = 0; c < nContacts; c++) unpackContact(contacts + 16 * c, w.contacts[c]);
  if (steps)
    for (size_t k = 0; k < w.params.size(); k++) w.params[k].step = steps[k];
}

/** The FP64 world takes the FP32 mirror's collision and topology products (the solve's inputs). */
void copyTopology(World<double> &d, const World<float> &f) {
  for (size_t i = 0; i < f.n(); i++) d.bodies[i].stamp = f.bodies[i].stamp;
  d.refs = f.refs;
  d.params = f.params;
  d.pairs = f.pairs;
  d.manifolds.resize(f.manifolds.size());
  for (size_t m = 0; m < f.manifolds.size(); m++) {
    const Manifold<float> &s = f.manifolds[m];
    Manifold<double> &o = d.manifolds[m];
    o.a = s.a; o.b = s.b; o.first = s.first; o.countw = s.countw;
    o.normal = V3<double>(s.normal.x, s.normal.y, s.normal.z);
    o.friction = s.friction;
  }
  d.contacts.resize(f.contacts.size());
  for (size_t c = 0; c < f.contacts.size(); c++) {
    const Contact<float> &s = f.contacts[c];
    Contact<double> &o = d.contacts[c];
    o.rA = V3<double>(s.rA.x, s.rA.y, s.rA.z); o.key = s.key;
    o.rB = V3<double>(s.rB.x, s.rB.y, s.rB.z); o.c0x = s.c0x;
    o.pen = V3<double>(s.pen.x, s.pen.y, s.pen.z); o.c0y = s.c0y;
    o.lam = V3<double>(s.lam.x, s.lam.y, s.lam.z); o.c0z = s.c0z;
  }
  d.adjStart = f.adjStart;
  d.adjList = f.adjList;
  d.colorA = f.colorA;
  d.colorStart = f.colorStart;
  d.colorBodies = f.colorBodies;
  d.clashes = f.clashes;
  d.overflow = f.overflow;
  d.numColors = f.numColors;
}

}  // namespace

extern "C" {

/**
 * worldParams: 16 floats per world: dt, gravity, betaLin, betaAng, gamma, alpha, up xyz, flags (bits), step (bits).
 * precision: 32 = FP32 mirror; 64 = FP64 (its step takes the collision/topology of an internal FP32 mirror).
 */
void *gs_create(int precision, uint32_t nBodies, const float *bodyRecs, const uint32_t *bodyWorld, uint32_t nWorlds,
                const float *worldParams, uint32_t nJoints, const float *jointRecs, const int32_t *info,
                uint32_t nNoCollide, const uint32_t *noCollide, const uint32_t *filters, uint32_t colorCap,
                uint32_t colorRounds, int iterations) {
  auto *h = new Handle();
  h->precision = precision;
  h->f = std::make_unique<World<float>>();
  build(*h->f, nBodies, bodyRecs, bodyWorld, nWorlds, worldParams, nJoints, jointRecs, info, nNoCollide, noCollide,
        filters);
  h->f->colorCap = colorCap;
  h->f->colorRounds = colorRounds;
  h->f->iterations = iterations;
  if (precision == 64) {
    h->d = std::make_unique<World<double>>();
    build(*h->d, nBodies, bodyRecs, bodyWorld, nWorlds, worldParams, nJoints, jointRecs, info, nNoCollide, noCollide,
          filters);
    h->d->colorCap = colorCap;
    h->d->colorRounds = colorRounds;
    h->d->iterations = iterations;
  }
  return h;
}

void gs_destroy(void *p) { delete static_cast<Handle *>(p); }

/** The frames constraints take at creation (rest orientations of angle-locked joints, motors' and limits' reference
 * tangents) from the current poses: for all such slots (n == 0) or the listed ones. */
void gs_capture_frames(void *p, const uint32_t *slots, uint32_t n) {
  auto *h = static_cast<Handle *>(p);
  if (n == 0) {
    h->f->captureFrames();
    if (h->d) h->d->captureFrames();
    return;
  }
  for (uint32_t e = 0; e < n; e++) {
    h->f->captureFrame(h->f->joints[slots[e]]);
    if (h->d) h->d->captureFrame(h->d->joints[slots[e]]);
  }
}

/** Constraint record edits: n (slot, word, value bits) triples; words 32..34 set the type, body A and body B. */
void gs_edit_joints(void *p, const uint32_t *triples, uint32_t n) {
  auto *h = static_cast<Handle *>(p);
  auto apply = [&](auto &w) {
    for (uint32_t e = 0; e < n; e++) {
      const uint32_t slot = triples[3 * e], word = triples[3 * e + 1], value = triples[3 * e + 2];
      auto &k = w.joints[slot];
      if (word == 32) k.type = int32_t(value);
      else if (word == 33) k.a = int32_t(value);
      else if (word == 34) k.b = int32_t(value);
      else *jointWord(k, word) = bitsToFloat(value);
    }
  };
  apply(*h->f);
  if (h->d) apply(*h->d);
}

/** Colour cap and Jones–Plassmann rounds from the next step on (the GPU world's setColoring / adapt). */
void gs_set_coloring(void *p, uint32_t colorCap, uint32_t colorRounds) {
  auto *h = static_cast<Handle *>(p);
  h->f->colorCap = colorCap;
  h->f->colorRounds = colorRounds;
  if (h->d) {
    h->d->colorCap = colorCap;
    h->d->colorRounds = colorRounds;
  }
}

/**
 * The first part of a step only, for checks of the warm start: stage 1 runs collision and topology (the contacts as
 * the narrowphase warm-started them, before any solve), stage 2 also the joint and body warm starts (each body's
 * initial guess). FP32 worlds only.
 */
void gs_partial_step(void *p, int stage) {
  auto *h = static_cast<Handle *>(p);
  h->f->collisionAndTopology();
  if (stage >= 2) {
    h->f->warmStartJoints();
    h->f->warmStartBodies();
  }
}

/** Pair capacity (0: unbounded): a step with more broadphase pairs overflows and is held, as on the GPU. */
void gs_set_pair_capacity(void *p, uint32_t capacity) { static_cast<Handle *>(p)->f->pairCapacity = capacity; }

/** 1 if the last step overflowed its pair capacity. */
uint32_t gs_overflow(void *p) { return static_cast<Handle *>(p)->f->overflow; }

void gs_set_brute_force(void *p, int on) {
  auto *h = static_cast<Handle *>(p);
  h->f->bruteForcePairs = on != 0;
  if (h->d) h->d->bruteForcePairs = on != 0;
}

/**
 * Seed the persistent state: reference poses (8 floats per body), colour state, the last step's manifolds (8 words)
 * and contacts (16 words) as the warm-start source, and each world's step counter.
 */
void gs_set_state(void *p, const float *refs, const uint32_t *colors, uint32_t nManifolds, const uint32_t *manifolds,
                  uint32_t nContacts, const uint32_t *contacts, const uint32_t *steps) {
  auto *h = static_cast<Handle *>(p);
  setState(*h->f, refs, colors, nManifolds, manifolds, nContacts, contacts, steps);
  if (h->d) setState(*h->d, refs, colors, nManifolds, manifolds, nContacts, contacts, steps);
}

void gs_step(void *p, int n) {
  auto *h = static_cast<Handle *>(p);
  for (int s = 0; s < n; s++) {
    if (h->precision == 64) {
      // Poses enter the FP32 mirror's collision as the FP64 world holds them, rounded to f32
      for (size_t i = 0; i < h->d->n(); i++) {
        float rec[40];
        packBody(h->d->bodies[i], rec);
        rec[27] = bitsToFloat(h->f->bodies[i].stamp);
        unpackBody(rec, h->f->bodies[i]);
      }
      h->f->contacts.resize(h->d->contacts.size());
      for (size_t c = 0; c < h->d->contacts.size(); c++) {
        const Contact<double> &s = h->d->contacts[c];
        Contact<float> &o = h->f->contacts[c];
        o.rA = V3<float>(float(s.rA.x), float(s.rA.y), float(s.rA.z)); o.key = s.key;
        o.rB = V3<float>(float(s.rB.x), float(s.rB.y), float(s.rB.z)); o.c0x = float(s.c0x);
        o.pen = V3<float>(float(s.pen.x), float(s.pen.y), float(s.pen.z)); o.c0y = float(s.c0y);
        o.lam = V3<float>(float(s.lam.x), float(s.lam.y), float(s.lam.z)); o.c0z = float(s.c0z);
      }
      for (size_t j = 0; j < h->d->joints.size(); j++) {
        float rec[32];
        packJoint(h->d->joints[j], rec);
        int32_t info[4] = {h->d->joints[j].type, h->d->joints[j].a, h->d->joints[j].b, 0};
        unpackJoint(rec, info, h->f->joints[j]);
      }
      const bool bounce = h->d->bouncing();
      if (bounce) h->d->restitutionBefore();
      h->f->collisionAndTopology();
      copyTopology(*h->d, *h->f);
      h->d->solve();
      if (bounce) h->d->restitutionAfter();
    } else {
      h->f->step();
    }
  }
}

/** counts: pairs, manifolds, contacts, clashes, colours used, bodies, joints. */
void gs_counts(void *p, uint32_t *out) {
  auto *h = static_cast<Handle *>(p);
  World<float> &w = *h->f;
  out[0] = uint32_t(w.pairs.size());
  out[1] = uint32_t(w.manifolds.size());
  out[2] = uint32_t(w.contacts.size());
  out[3] = w.clashes;
  out[4] = w.numColors;
  out[5] = uint32_t(w.n());
  out[6] = uint32_t(w.joints.size());
  out[7] = w.pivotFallbacks;
}

/** Bodies as 40 values each (inertialPos.w carries the moved-step stamp as its integer value). */
void gs_read_bodies(void *p, double *out) {
  auto *h = static_cast<Handle *>(p);
  if (h->d)
    for (size_t i = 0; i < h->d->n(); i++) packBody(h->d->bodies[i], out + 40 * i);
  else
    for (size_t i = 0; i < h->f->n(); i++) {
      packBody(h->f->bodies[i], out + 40 * i);
      out[40 * i + 27] = double(h->f->bodies[i].stamp);
    }
}
void gs_read_joints(void *p, double *out) {
  auto *h = static_cast<Handle *>(p);
  if (h->d)
    for (size_t j = 0; j < h->d->joints.size(); j++) packJoint(h->d->joints[j], out + 32 * j);
  else
    for (size_t j = 0; j < h->f->joints.size(); j++) packJoint(h->f->joints[j], out + 32 * j);
}
void gs_read_pairs(void *p, uint32_t *out) {
  auto *h = static_cast<Handle *>(p);
  for (size_t k = 0; k < h->f->pairs.size(); k++) {
    out[2 * k] = h->f->pairs[k].first;
    out[2 * k + 1] = h->f->pairs[k].second;
  }
}
/** Manifolds in the 8-word record (first = index into the contact list below). */
void gs_read_manifolds(void *p, uint32_t *out) {
  auto *h = static_cast<Handle *>(p);
  for (size_t m = 0; m < h->f->manifolds.size(); m++) packManifold(h->f->manifolds[m], out + 8 * m);
}
/** Contacts: keys (u32 per contact) and 16 values each (key slot 3 left 0). FP64 worlds give their own values. */
void gs_read_contacts(void *p, uint32_t *keys, double *out) {
  auto *h = static_cast<Handle *>(p);
  for (size_t c = 0; c < h->f->contacts.size(); c++) {
    if (keys) keys[c] = h->d ? h->d->contacts[c].key : h->f->contacts[c].key;
    if (out) {
      if (h->d) packContact(h->d->contacts[c], out + 16 * c);
      else packContact(h->f->contacts[c], out + 16 * c);
    }
  }
}
/** Raw contact records (16 words, FP32 mirror) — for seeding another run. */
void gs_read_contact_words(void *p, uint32_t *out) {
  auto *h = static_cast<Handle *>(p);
  for (size_t c = 0; c < h->f->contacts.size(); c++) {
    const Contact<float> &k = h->f->contacts[c];
    const float v[16] = {k.rA.x, k.rA.y, k.rA.z, 0, k.rB.x, k.rB.y, k.rB.z, k.c0x,
                         k.pen.x, k.pen.y, k.pen.z, k.c0y, k.lam.x, k.lam.y, k.lam.z, k.c0z};
    for (int i = 0; i < 16; i++) out[16 * c + i] = floatToBits(v[i]);
    out[16 * c + 3] = k.key;
  }
}
void gs_read_colors(void *p, uint32_t *out) {
  auto *h = static_cast<Handle *>(p);
  std::memcpy(out, h->f->colorA.data(), 4 * h->f->n());
}
/** Adjacency: start (bodies + 1) then the list. */
uint32_t gs_read_adjacency(void *p, uint32_t *start, uint32_t *list) {
  auto *h = static_cast<Handle *>(p);
  if (start) std::memcpy(start, h->f->adjStart.data(), 4 * h->f->adjStart.size());
  if (list) std::memcpy(list, h->f->adjList.data(), 4 * h->f->adjList.size());
  return uint32_t(h->f->adjList.size());
}
void gs_read_refs(void *p, float *out) {
  auto *h = static_cast<Handle *>(p);
  for (size_t i = 0; i < h->f->n(); i++) {
    std::memcpy(out + 8 * i, h->f->refs[i].pos, 16);
    std::memcpy(out + 8 * i + 4, h->f->refs[i].rot, 16);
  }
}

// --- edits between steps ---------------------------------------------------------------------------------------------

/** Body commands (12 words each, World::applyCommands), in submission order. */
void gs_apply_commands(void *p, const uint32_t *words, uint32_t n) {
  auto *h = static_cast<Handle *>(p);
  if (h->f->sleepOn) h->f->sleepWakeAll();  // body commands wake every body first (prepareBodyCommands)
  h->f->applyCommands(words, n);
  if (h->d) h->d->applyCommands(words, n);
}

/** World k's parameters from 10 words: dt, gravity, betaLin, betaAng, gamma, alpha, up xyz, flags (bits); the step
 * counter is kept. */
void gs_set_params(void *p, uint32_t k, const float *v) {
  auto *h = static_cast<Handle *>(p);
  for (int pass = 0; pass < 2; pass++) {
    std::vector<WorldParams> &params = pass == 0 ? h->f->params : h->d ? h->d->params : h->f->params;
    if (pass == 1 && !h->d) break;
    WorldParams &q = params[k];
    q.dt = v[0]; q.gravity = v[1]; q.betaLin = v[2]; q.betaAng = v[3]; q.gamma = v[4]; q.alpha = v[5];
    q.up[0] = v[6]; q.up[1] = v[7]; q.up[2] = v[8];
    q.flags = floatToBits(v[9]);
  }
}

/** Threads of the colour-parallel variant (OpenMP); 1 (the default) runs every loop in order. Results are the same. */
void gs_set_threads(int threads) { omp_set_num_threads(threads < 1 ? 1 : threads); }
int gs_get_threads() { return omp_get_max_threads(); }

void gs_set_iterations(void *p, int iterations) {
  auto *h = static_cast<Handle *>(p);
  h->f->iterations = iterations;
  if (h->d) h->d->iterations = iterations;
}

/** Collision filters (groups, collides-with) of n bodies. */
void gs_set_filters(void *p, const uint32_t *indices, const uint32_t *groups, const uint32_t *masks, uint32_t n) {
  auto *h = static_cast<Handle *>(p);
  if (h->f->sleepOn) h->f->sleepWakeAll();
  for (uint32_t k = 0; k < n; k++) {
    h->f->filters[2 * size_t(indices[k])] = groups[k];
    h->f->filters[2 * size_t(indices[k]) + 1] = masks[k];
    if (h->d) {
      h->d->filters[2 * size_t(indices[k])] = groups[k];
      h->d->filters[2 * size_t(indices[k]) + 1] = masks[k];
    }
  }
}

/** Restitution coefficients of n bodies (the bounce passes run while any is above 0). */
void gs_set_restitution(void *p, const uint32_t *indices, const float *values, uint32_t n) {
  auto *h = static_cast<Handle *>(p);
  if (h->f->sleepOn) h->f->sleepWakeAll();
  for (uint32_t k = 0; k < n; k++) {
    h->f->restitution[indices[k]] = values[k];
    if (h->d) h->d->restitution[indices[k]] = double(values[k]);
  }
}

// --- sleeping (FP32 mirror only) ------------------------------------------------------------------------------------

/** Enable sleeping (eligible: one byte per body, null = all); returns -1 for an FP64 handle. */
int gs_enable_sleeping(void *p, float speed, float delay, int islands, int startAsleep, const uint8_t *eligible) {
  auto *h = static_cast<Handle *>(p);
  if (h->d) return -1;
  h->f->sleepEnable(speed, delay, islands != 0, startAsleep != 0, eligible);
  return 0;
}
void gs_disable_sleeping(void *p) { static_cast<Handle *>(p)->f->sleepDisable(); }
void gs_wake_all(void *p) {
  auto *h = static_cast<Handle *>(p);
  if (h->f->sleepOn) h->f->sleepWakeAll();
}
/** wakeBody() for n bodies: held awake through the next step; in island mode their islands wake. */
void gs_wake_bodies(void *p, const uint32_t *indices, uint32_t n) {
  auto *h = static_cast<Handle *>(p);
  if (h->f->sleepOn) h->f->sleepHold(indices, n);
}
/** Sleep eligibility of n bodies, then every body wakes (setSleepEnabled). */
void gs_set_sleep_enabled(void *p, const uint32_t *indices, const uint8_t *flags, uint32_t n) {
  auto *h = static_cast<Handle *>(p);
  if (!h->f->sleepOn) return;
  for (uint32_t k = 0; k < n; k++) h->f->sleep[indices[k]].eligible = flags[k];
  h->f->sleepWakeAll();
}
/** World k's sleeping count and wake-all request after the last step. */
void gs_sleep_stats(void *p, uint32_t k, uint32_t *out) {
  auto *h = static_cast<Handle *>(p);
  out[0] = h->f->sleepOn ? h->f->sleepGlobals[4 * k + 1] : 0u;
  out[1] = h->f->sleepOn ? h->f->sleepGlobals[4 * k] : 0u;
}
/** Every body's sleep state, 8 words: mass (f32 bits), quiet (f32 bits), asleep, support, eligible, parent, wake,
 * flags. */
void gs_read_sleep(void *p, uint32_t *out) {
  auto *h = static_cast<Handle *>(p);
  for (size_t i = 0; i < h->f->sleep.size(); i++) {
    const auto &s = h->f->sleep[i];
    uint32_t *o = out + 8 * i;
    o[0] = floatToBits(s.mass);
    o[1] = floatToBits(s.quiet);
    o[2] = s.asleep; o[3] = s.support; o[4] = s.eligible; o[5] = s.parent; o[6] = s.wake; o[7] = s.flags;
  }
}

/** Seed every body's sleep state (gs_read_sleep's layout) and every world's 4 sleep globals. */
void gs_set_sleep(void *p, const uint32_t *states, const uint32_t *globals) {
  auto *h = static_cast<Handle *>(p);
  if (!h->f->sleepOn) return;
  for (size_t i = 0; i < h->f->sleep.size(); i++) {
    auto &s = h->f->sleep[i];
    const uint32_t *o = states + 8 * i;
    s.mass = bitsToFloat(o[0]);
    s.quiet = bitsToFloat(o[1]);
    s.asleep = o[2]; s.support = o[3]; s.eligible = o[4]; s.parent = o[5]; s.wake = o[6]; s.flags = o[7];
  }
  std::copy(globals, globals + h->f->sleepGlobals.size(), h->f->sleepGlobals.begin());
}

// --- shape queries ---------------------------------------------------------------------------------------------------

/**
 * raycastAll / sphereCastAll over n body records (40 floats) with their filters (group, collides-with): queries of 8
 * floats (origin xyz, max distance, unit direction xyz, radius); hits of 8 floats (normal xyz, distance or -1, index
 * bits, 0, 0, 0). A body is left out when group & mask is 0, when (with useMembership) membership is 0 or its
 * collides-with & membership is 0, or when it is in `ignore`.
 */
void gs_cast(uint32_t nBodies, const float *bodyRecs, const uint32_t *filters, const float *queries, uint32_t n,
             const uint32_t *ignore, uint32_t nIgnore, uint32_t mask, uint32_t membership, int useMembership,
             float *hits, const uint32_t *hullWords) {
  for (uint32_t q = 0; q < n; q++) {
    const float *query = queries + 8ull * q;
    float bestT = INFINITY;
    uint32_t bestI = 0xffffffffu;
    oracle_queries::Hit best{-1.f, {0, 0, 0}};
    for (uint32_t i = 0; i < nBodies; i++) {
      if ((filters[2 * i] & mask) == 0u) continue;
      if (useMembership && (membership == 0u || (filters[2 * i + 1] & membership) == 0u)) continue;
      bool skip = false;
      for (uint32_t k = 0; k < nIgnore; k++) skip = skip || ignore[k] == i;
      if (skip) continue;
      const oracle_queries::Hit h = oracle_queries::trace(bodyRecs + 40ull * i, query, hullWords);
      if (h.t < 0.f) continue;
      const float t = h.t > 0.f ? h.t : 0.f;
      if (t < bestT) {  // bodies in index order: a tie keeps the smaller index
        bestT = t;
        bestI = i;
        best = h;
      }
    }
    float *o = hits + 8ull * q;
    for (int k = 0; k < 8; k++) o[k] = 0.f;
    uint32_t idx = bestI;
    std::memcpy(o + 4, &idx, 4);
    if (bestI == 0xffffffffu) {
      o[3] = -1.f;
      continue;
    }
    o[0] = best.n.x;
    o[1] = best.n.y;
    o[2] = best.n.z;
    o[3] = bestT;
  }
}

// --- unit-level probes (for tests of single functions against the upstream restatement) -----------------------------

/**
 * Box/box contacts in the GPU semantics (faceBias flag), for comparison with the upstream collide(): returns the
 * contact count and writes feature keys, world points on A and B (3 + 3 doubles each) and the normal (B to A).
 */
int gs_probe_box_contacts(int precision, const double *poseA, const double *sizeA, const double *poseB,
                          const double *sizeB, int faceBias, int maxPoly, uint32_t *features, double *points,
                          double *normal) {
  auto run = [&](auto zero) {
    using R = decltype(zero);
    Box<R> A = makeBox(V3<R>(R(poseA[0]), R(poseA[1]), R(poseA[2])),
                       Q4<R>(R(poseA[3]), R(poseA[4]), R(poseA[5]), R(poseA[6])),
                       V3<R>(R(sizeA[0]), R(sizeA[1]), R(sizeA[2])));
    Box<R> B = makeBox(V3<R>(R(poseB[0]), R(poseB[1]), R(poseB[2])),
                       Q4<R>(R(poseB[3]), R(poseB[4]), R(poseB[5]), R(poseB[6])),
                       V3<R>(R(sizeB[0]), R(sizeB[1]), R(sizeB[2])));
    Sat<R> sat;
    Found<R> f = collideBoxes(A, B, faceBias != 0, maxPoly, sat);
    for (int i = 0; i < f.count; i++) {
      features[i] = f.feature[i];
      double v[6] = {double(f.xA[i].x), double(f.xA[i].y), double(f.xA[i].z),
                     double(f.xB[i].x), double(f.xB[i].y), double(f.xB[i].z)};
      std::memcpy(points + 6 * i, v, sizeof v);
    }
    normal[0] = -double(sat.n.x);
    normal[1] = -double(sat.n.y);
    normal[2] = -double(sat.n.z);
    return f.count;
  };
  return precision == 64 ? run(0.0) : run(0.0f);
}

/**
 * Contacts of one pair of body records (40 f32 each) by shape — box/box, sphere, capsule — as the narrowphase makes
 * them: features, world points on A and B (6 per point) and the manifold normal (from B towards A). Returns the count.
 */
int gs_probe_pair_contacts(int precision, const float *recA, const float *recB, int faceBias, uint32_t *features,
                           double *points, double *normal) {
  auto run = [&](auto zero) {
    using R = decltype(zero);
    Body<R> BA, BB;
    unpackBody(recA, BA);
    unpackBody(recB, BB);
    Box<R> A = makeBox(BA.pos, BA.rot, BA.size), B = makeBox(BB.pos, BB.rot, BB.size);
    Sat<R> sat;
    Found<R> f = shapeContacts(BA, BB, A, B, faceBias != 0, sat);
    for (int i = 0; i < f.count; i++) {
      features[i] = f.feature[i];
      double v[6] = {double(f.xA[i].x), double(f.xA[i].y), double(f.xA[i].z),
                     double(f.xB[i].x), double(f.xB[i].y), double(f.xB[i].z)};
      std::memcpy(points + 6 * i, v, sizeof v);
    }
    normal[0] = -double(sat.n.x);
    normal[1] = -double(sat.n.y);
    normal[2] = -double(sat.n.z);
    return f.count;
  };
  return precision == 64 ? run(0.0) : run(0.0f);
}

/** The hull buffer (4 u32 per word) every hull body's shape code indexes. */
void gs_set_hulls(void *p, const uint32_t *words, uint32_t n) {
  auto *h = static_cast<Handle *>(p);
  h->f->hulls.w.assign(words, words + 4ull * n);
  if (h->d) h->d->hulls.w.assign(words, words + 4ull * n);
}

/**
 * Contacts of one pair of body records with a hull buffer (any shapes), as gs_probe_pair_contacts; for two boxes or
 * hulls also the three separating-axis queries of collidePoly in `queries` (face of A: sep, index; face of B: sep,
 * index; best edge pair: sep, valid, A edge, B edge), each -3.4e38 where the query did not run.
 */
int gs_probe_hull_pair(int precision, const float *recA, const float *recB, const uint32_t *words, uint32_t nWords,
                       int faceBias, uint32_t *features, double *points, double *normal, double *queries) {
  HullBuffer H;
  H.w.assign(words, words + 4ull * nWords);
  auto run = [&](auto zero) {
    using R = decltype(zero);
    Body<R> BA, BB;
    unpackBody(recA, BA);
    unpackBody(recB, BB);
    Box<R> A = makeBox(BA.pos, BA.rot, BA.size), B = makeBox(BB.pos, BB.rot, BB.size);
    Sat<R> sat;
    Found<R> f;
    for (int k = 0; k < 8; k++) queries[k] = -3.4e38;
    const bool polyPair = BA.shape != R(SHAPE_SPHERE) && BB.shape != R(SHAPE_SPHERE) && BA.shape != R(SHAPE_CAPSULE) &&
                          BB.shape != R(SHAPE_CAPSULE) && (BA.shape >= R(SHAPE_HULL_F) || BB.shape >= R(SHAPE_HULL_F));
    if (polyPair) {
      PolyQueries<R> q;
      q.fa.sep = q.fb.sep = R(-3.4e38f);
      q.fa.face = q.fb.face = 0;
      f = collidePoly(H, polyOf(H, BA.pos, BA.rot, BA.size, BA.shape), polyOf(H, BB.pos, BB.rot, BB.size, BB.shape),
                      faceBias != 0, sat, &q);
      queries[0] = double(q.fa.sep);
      queries[1] = q.fa.face;
      queries[2] = double(q.fb.sep);
      queries[3] = q.fb.face;
      queries[4] = double(q.eq.sep);
      queries[5] = q.eq.valid ? 1 : 0;
      queries[6] = q.eq.ea;
      queries[7] = q.eq.eb;
    } else {
      f = shapeContacts(BA, BB, A, B, faceBias != 0, sat, &H);
    }
    for (int i = 0; i < f.count; i++) {
      features[i] = f.feature[i];
      double v[6] = {double(f.xA[i].x), double(f.xA[i].y), double(f.xA[i].z),
                     double(f.xB[i].x), double(f.xB[i].y), double(f.xB[i].z)};
      std::memcpy(points + 6 * i, v, sizeof v);
    }
    normal[0] = -double(sat.n.x);
    normal[1] = -double(sat.n.y);
    normal[2] = -double(sat.n.z);
    return f.count;
  };
  return precision == 64 ? run(0.0) : run(0.0f);
}

/** LDLᵀ 6x6 solve (lower triangle of the symmetric 6x6 `a`, row-major 36) with an optional pivot floor. */
void gs_probe_solve6(const double *a, const double *b, double pivotFloor, double *x) {
  M3<double> lin, ang, crs;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      lin.m[r * 3 + c] = a[r * 6 + c];
      ang.m[r * 3 + c] = a[(r + 3) * 6 + c + 3];
      crs.m[r * 3 + c] = a[(r + 3) * 6 + c];
    }
  V3<double> xl, xa;
  solve6(lin, ang, crs, V3<double>(b[0], b[1], b[2]), V3<double>(b[3], b[4], b[5]), xl, xa, pivotFloor);
  x[0] = xl.x; x[1] = xl.y; x[2] = xl.z; x[3] = xa.x; x[4] = xa.y; x[5] = xa.z;
}

/** Priority (lowbias32 of index + 0x9e3779b9) used by the colouring. */
uint32_t gs_probe_priority(uint32_t i) { return hash32(i + 0x9e3779b9u); }
}

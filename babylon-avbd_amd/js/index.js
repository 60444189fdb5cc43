'use strict';
// JavaScript world API over the N-API addon (avbd_napi.node → libavbd_hip.so): the reference's
// createSolverPipeline({backend}) (src/native/pipeline.js:14-93) with a "hip" backend, its AvbdScene3D builder
// (src/native/scenes.js:107-702, bodies and joints) and a world with the shape of its GPU worlds
// (src/native/solver3D.js:104-368; scenes.d.ts:317-454): step, advance, reads returning Promises, adapt, destroy.
// Written for Node >= 12 (CommonJS, no optional chaining). Records are packed exactly as the Python builder of this
// package packs them, so one scene gives the same bytes from either host.

const path = require('path');

let addon = null;
function native() {
  if (!addon) addon = require(path.join(__dirname, 'avbd_napi.node'));
  return addon;
}

const BIG = 3.0e38;
const T_JOINT = 1;
const T_SPRING = 2;
const SHAPE_BOX = 0;
const SHAPE_SPHERE = 1;
const SHAPE_CAPSULE = -1;
const SHAPE_HULL = 3; // + the word of the body's hull record in the scene's hull buffer
const FLAGS_DEFAULT = 31;
const MAX_HULL_VERTICES = 32;

function finite(x) {
  if (x === Infinity) return BIG;
  if (x === -Infinity) return -BIG;
  return x;
}

function vector(v, n, name) {
  if (!v || v.length !== n) throw Error(`${name} requires ${n} finite components`);
  const a = Array.from(v, Number);
  if (!a.every(Number.isFinite)) throw Error(`${name} requires ${n} finite components`);
  return a;
}

function positive(v, name, zero) {
  const x = Number(v);
  if (!Number.isFinite(x) || (zero ? x < 0 : x <= 0))
    throw Error(`${name} must be ${zero ? 'nonnegative' : 'positive'} and finite`);
  return x;
}

/** ref/math.ts rotate (x, y, z, w quaternion). */
function rotate(q, v) {
  const ux = q[0], uy = q[1], uz = q[2], w = q[3];
  const tx = (uy * v[2] - uz * v[1]) * 2;
  const ty = (uz * v[0] - ux * v[2]) * 2;
  const tz = (ux * v[1] - uy * v[0]) * 2;
  return [v[0] + tx * w + (uy * tz - uz * ty), v[1] + ty * w + (uz * tx - ux * tz), v[2] + tz * w + (ux * ty - uy * tx)];
}

/** Box mass, principal moments and bounding radius (ref/body.ts:38-47). */
function boxMass(s, density) {
  const m = s[0] * s[1] * s[2] * density;
  const moment = [((s[1] * s[1] + s[2] * s[2]) / 12) * m, ((s[0] * s[0] + s[2] * s[2]) / 12) * m,
    ((s[0] * s[0] + s[1] * s[1]) / 12) * m];
  const hx = s[0] * 0.5, hy = s[1] * 0.5, hz = s[2] * 0.5;
  return [m, moment, Math.sqrt(hx * hx + hy * hy + hz * hz)];
}

/** Solid capsule along local Y of full tip-to-tip height (src/gpu/capsuleShape.js:28-38). */
function capsuleMass(radius, height, density) {
  const length = height - 2 * radius;
  const r2 = radius * radius, l2 = length * length;
  const cylinder = Math.PI * radius * radius * length * density;
  const caps = (4 / 3) * Math.PI * (r2 * radius) * density;
  const axial = (cylinder * r2) / 2 + (caps * 2 * r2) / 5;
  const radial = cylinder * (l2 / 12 + r2 / 4) + caps * ((2 * r2) / 5 + l2 / 4 + (3 * length * radius) / 8);
  return [cylinder + caps, [radial, axial, radial], height / 2];
}

/** One 40-float body record as GpuSolver3D.pack writes a freshly built body (solver.ts:794-820). */
function bodyRecord(position, size, mass, moment, radius, friction, shape) {
  const r = new Float32Array(40);
  r.set(position, 0);
  r[3] = friction;
  r[7] = 1; // rotation (0, 0, 0, 1)
  r[15] = 1; // initialRot
  r.set(size, 16);
  r[19] = mass;
  r.set(moment, 20);
  r[23] = radius;
  r[31] = 1; // inertialRot
  r[39] = shape;
  return r;
}

/**
 * A cooked convex hull (reference/three-avbd/src/avbd3d/hull.ts HullShape): the packed record the narrowphase reads
 * (header at word 0; `record(at)` rebases it), mass properties per unit density and the principal frame, with the
 * vertices, faces and edges decoded from the record.
 */
class HullShape {
  constructor(cooked) {
    this.words = cooked.record.length / 4;
    this.data = cooked.record;
    Object.assign(this, { volume: cooked.volume, moments: cooked.moments, size: cooked.size, radius: cooked.radius,
      center: cooked.center, rotation: cooked.rotation });
    const u = this.data, f = new Float32Array(u.buffer, u.byteOffset, u.length);
    const [vs, nv, fs, nf] = [u[0], u[1], u[2], u[3]];
    const [es, ne, is] = [u[4], u[5], u[6]];
    this.vertices = new Float32Array(3 * nv);
    for (let v = 0; v < nv; v++) this.vertices.set(f.subarray(4 * (vs + v), 4 * (vs + v) + 3), 3 * v);
    this.faces = [];
    for (let k = 0; k < nf; k++) {
      const w = 4 * (fs + 2 * k), first = u[w + 4], count = u[w + 5];
      this.faces.push({ normal: Array.from(f.subarray(w, w + 3)), d: f[w + 3],
        verts: Array.from(u.subarray(4 * is + first, 4 * is + first + count)) });
    }
    this.edges = [];
    for (let k = 0; k < ne; k++) this.edges.push(Array.from(u.subarray(4 * (es + k), 4 * (es + k) + 4)));
  }

  record(at) {
    const r = Uint32Array.from(this.data);
    for (const w of [0, 2, 4, 6]) r[w] += at;
    return r;
  }
}

function flatPoints(points, name) {
  const a = Array.from(points, Number);
  if (a.length % 3 || !a.every(Number.isFinite)) throw Error(`${name} must contain finite xyz triples`);
  return Float64Array.from(a);
}

/** The convex hull of xyz points (at most MAX_HULL_VERTICES vertices), or null when they span no volume (hull.ts:60). */
function convexHull(points) {
  const cooked = native().hullFromPoints(flatPoints(points, 'Hull points'));
  return cooked ? new HullShape(cooked) : null;
}

/** A hull from a closed outward triangulation of a convex point set, or null (hull.ts:208). */
function hullFromTriangles(positions, indices) {
  const idx = Array.from(indices, Number);
  const n = positions.length / 3;
  if (idx.length % 3 || !idx.every((i) => Number.isInteger(i) && i >= 0 && i < n)) return null;
  const cooked = native().hullFromTriangles(flatPoints(positions, 'positions'), Uint32Array.from(idx));
  return cooked ? new HullShape(cooked) : null;
}

// --- terrain (src/native/terrain.js): every triangle a static prism hull ---------------------------------------------

// The closed prism over top corners 0, 1, 2 and their base copies 3, 4, 5, and the same with each triangle turned
const PRISM_FACES = [0, 1, 2, 3, 5, 4, 0, 3, 4, 0, 4, 1, 1, 4, 5, 1, 5, 2, 2, 5, 3, 2, 3, 0];
const REVERSED_PRISM_FACES = PRISM_FACES.map((v, i) => PRISM_FACES[i - (i % 3) + [0, 2, 1][i % 3]]);

function finiteList(values, length, name) {
  const a = Array.from(values || [], Number);
  if (a.length !== length || !a.every(Number.isFinite)) throw Error(`${name} requires ${length} finite values`);
  return a;
}

function terrainPiece(points) {
  const center = [0, 1, 2].map((k) => points.reduce((sum, p) => sum + p[k], 0) / points.length);
  const [a, b, c, back] = points;
  const u = b.map((v, k) => v - a[k]), v = c.map((x, k) => x - a[k]);
  const outward = (u[1] * v[2] - u[2] * v[1]) * (a[0] - back[0]) + (u[2] * v[0] - u[0] * v[2]) * (a[1] - back[1]) +
    (u[0] * v[1] - u[1] * v[0]) * (a[2] - back[2]);
  const shape = hullFromTriangles([].concat(...points.map((p) => p.map((x, k) => x - center[k]))),
    outward > 0 ? PRISM_FACES : REVERSED_PRISM_FACES);
  if (!shape) throw Error('Terrain contains a degenerate collision piece');
  return { shape, position: shape.center.map((x, k) => x + center[k]), rotation: shape.rotation };
}

/** Static solid ground; heights are row-major, each cell keeps both triangles (terrain.js:62-134). */
function cookHeightfield(options) {
  const o = options || {};
  const { columns, rows } = o;
  for (const [name, value] of [['columns', columns], ['rows', rows]])
    if (!Number.isInteger(value) || value < 2) throw Error(`${name} must be an integer >= 2`);
  const values = finiteList(o.heights, columns * rows, 'heights');
  const offset = finiteList(o.origin || [0, 0, 0], 3, 'origin');
  const upAxis = o.upAxis === undefined ? 1 : o.upAxis;
  if (![0, 1, 2].includes(upAxis)) throw Error('upAxis must be 0, 1 or 2');
  const spacing = o.spacing === undefined ? 1 : o.spacing;
  const step = typeof spacing === 'number' ? [spacing, spacing] : finiteList(spacing, 2, 'spacing');
  step.forEach((v) => positive(v, 'spacing'));
  const base = o.baseHeight === undefined ? Math.min(...values) - 2 : o.baseHeight;
  if (!Number.isFinite(base) || values.some((h) => h <= base)) throw Error('baseHeight must be below every terrain height');
  const cells = (columns - 1) * (rows - 1);
  if (o.holes && o.holes.length !== cells) throw Error('holes requires one value per terrain cell');
  const axes = [0, 1, 2].filter((k) => k !== upAxis), pieces = [];
  const point = (x, z) => {
    const p = offset.slice();
    p[axes[0]] += x * step[0];
    p[axes[1]] += z * step[1];
    p[upAxis] += values[z * columns + x];
    return p;
  };
  for (let z = 0; z < rows - 1; z++)
    for (let x = 0; x < columns - 1; x++) {
      if (o.holes && o.holes[z * (columns - 1) + x]) continue;
      const a = point(x, z), b = point(x + 1, z), c = point(x, z + 1), d = point(x + 1, z + 1);
      for (const top of [[a, b, c], [b, d, c]])
        pieces.push(terrainPiece(top.concat(top.map((p) => { const q = p.slice(); q[upAxis] = offset[upAxis] + base; return q; }))));
    }
  return { kind: 'heightfield', pieces, triangleCount: pieces.length };
}

/** Triangle surfaces, each backed `thickness` against its winding normal (terrain.js:137-176). */
function cookStaticMesh(options) {
  const o = options || {};
  if (!o.positions || o.positions.length % 3) throw Error('positions requires xyz triples');
  const p = finiteList(o.positions, o.positions.length, 'positions'), count = p.length / 3;
  if (o.indices === undefined && count % 3) throw Error('Nonindexed positions require three vertices per triangle');
  const ids = o.indices === undefined ? Array.from({ length: count }, (_, i) => i) : Array.from(o.indices);
  if (!ids.length || ids.length % 3 || !ids.every((i) => Number.isInteger(i) && i >= 0 && i < count))
    throw Error('indices requires valid triangle vertex indices');
  const thickness = o.thickness === undefined ? 0.1 : o.thickness;
  positive(thickness, 'thickness');
  const pieces = [];
  for (let i = 0; i < ids.length; i += 3) {
    const top = ids.slice(i, i + 3).map((id) => p.slice(3 * id, 3 * id + 3)), [a, b, c] = top;
    const u = b.map((v, k) => v - a[k]), v = c.map((x, k) => x - a[k]);
    const n = [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]];
    const length = Math.hypot(...n);
    if (length === 0) continue; // imported meshes often carry empty seam triangles
    pieces.push(terrainPiece(top.concat(top.map((q) => q.map((x, k) => x - (n[k] * thickness) / length)))));
  }
  if (!pieces.length) throw Error('Static mesh requires at least one nondegenerate triangle');
  return { kind: 'static-mesh', pieces, triangleCount: pieces.length };
}

/** Static hull bodies for cooked pieces (terrain.js:179-218). */
function createTerrainBodies(scene, cooked, options) {
  const o = options || {};
  if (o.mass !== undefined && o.mass !== 0) throw Error('Terrain is static and requires zero mass');
  const friction = o.friction === undefined ? 0.6 : o.friction, restitution = o.restitution === undefined ? 0 : o.restitution;
  if (!Number.isFinite(friction) || friction < 0) throw Error('friction must be nonnegative and finite');
  if (!Number.isFinite(restitution) || restitution < 0 || restitution > 1) throw Error('restitution must be between 0 and 1');
  for (const name of ['group', 'collidesWith']) {
    const value = o[name] === undefined ? 0xffffffff : o[name];
    if (!Number.isInteger(value) || value < 0 || value > 0xffffffff) throw Error(`${name} must be an unsigned 32-bit integer`);
  }
  if (!cooked || !cooked.pieces || !cooked.pieces.length) throw Error('Terrain requires collision pieces');
  return cooked.pieces.map((p) => scene.addHullShape(p.shape, { density: 0, friction, position: p.position,
    rotation: p.rotation, allowSleep: false, restitution, group: o.group === undefined ? 0xffffffff : o.group,
    collidesWith: o.collidesWith === undefined ? 0xffffffff : o.collidesWith }));
}

/** Portable 3D scene builder (scenes.js:107-273): Y up, gravity -10, dt 1/60, 10 iterations by default. */
class AvbdScene3D {
  constructor(options) {
    const o = options || {};
    const gravity = o.gravity === undefined ? -10 : Number(o.gravity);
    if (!Number.isFinite(gravity)) throw Error('gravity must be finite');
    const iterations = o.iterations === undefined ? 10 : o.iterations;
    if (!Number.isInteger(iterations) || iterations < 1) throw Error('iterations must be an integer');
    this.params = {
      dt: positive(o.timeStep === undefined ? 1 / 60 : o.timeStep, 'timeStep'),
      gravity,
      iterations,
      alpha: o.alpha === undefined ? 0.99 : Number(o.alpha),
      gamma: o.gamma === undefined ? 0.999 : Number(o.gamma),
      betaLin: o.betaLin === undefined ? 1e4 : Number(o.betaLin),
      betaAng: o.betaAng === undefined ? 100 : Number(o.betaAng),
      flags: o.flags === undefined ? FLAGS_DEFAULT : o.flags >>> 0,
    };
    this.up = vector(o.up || [0, 1, 0], 3, 'up');
    if (Math.abs(Math.hypot(this.up[0], this.up[1], this.up[2]) - 1) > 1e-5) throw Error('up must be a unit vector');
    this.records = [];
    this.filters = new Map();
    this.restitution = new Map();  // body -> restitution coefficient above 0
    this.noSleep = new Set();      // bodies built with allowSleep false
    this.sleepOptIn = false;       // a body was built with allowSleep true: its world sleeps (scenes.js:233-240, 689)
    this.joints = [];
    this.ignore = [];
    this.hulls = new Map();  // HullShape -> header word in the scene's hull buffer (one record per shape)
    this.hullWords = 0;
    this.created = false;
  }

  get bodyCount() { return this.records.length; }

  /** scene.createSolver(pipeline, options) (scenes.js:652-702): a world of this scene on the pipeline's device. */
  createSolver(pipeline, options) {
    if (!pipeline || typeof pipeline.createSolver !== 'function') throw Error('A solver pipeline is required');
    return pipeline.createSolver(this, options);
  }

  assertEditable() {
    if (this.created) throw Error('Build this scene before createSolver; use the solver\'s edit methods afterwards');
  }

  add(rec, options) {
    this.assertEditable();
    const o = options || {};
    if (o.rotation !== undefined && o.rotation !== null) {
      const q = vector(o.rotation, 4, 'rotation');
      const l = Math.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
      if (!l) throw Error('rotation must be nonzero');
      rec.set(q.map((x) => x / l), 4);
    }
    if (o.mass !== undefined && o.mass !== null) {
      const mass = positive(o.mass, 'mass', true);
      if (rec[19] === 0 && mass > 0) throw Error('A positive mass needs positive density');
      const ratio = rec[19] > 0 ? mass / rec[19] : 0;
      const f = Math.fround(ratio);
      rec[19] = mass;
      for (let k = 20; k < 23; k++) rec[k] = Math.fround(rec[k] * f);
    }
    if (o.velocity !== undefined && o.velocity !== null) rec.set(vector(o.velocity, 3, 'velocity'), 32);
    if (o.angularVelocity !== undefined && o.angularVelocity !== null)
      rec.set(vector(o.angularVelocity, 3, 'angularVelocity'), 36);
    const restitution = o.restitution === undefined ? undefined : coefficient(o.restitution);
    const group = o.group === undefined ? 0xffffffff : o.group;
    const mask = o.collidesWith === undefined ? 0xffffffff : o.collidesWith;
    for (const m of [group, mask])
      if (!Number.isInteger(m) || m < 0 || m > 0xffffffff) throw Error('Collision masks must be unsigned 32-bit integers');
    if (o.allowSleep !== undefined && typeof o.allowSleep !== 'boolean') throw Error('allowSleep must be a boolean');
    const index = this.records.length;
    if (o.allowSleep === false) this.noSleep.add(index);
    this.sleepOptIn = this.sleepOptIn || o.allowSleep === true;
    if (group !== 0xffffffff || mask !== 0xffffffff) this.filters.set(index, [group, mask]);
    if (restitution) this.restitution.set(index, restitution);
    this.records.push(rec);
    return index;
  }

  /** A centred box with full XYZ extents; zero density makes it static (scenes.js:131-144). */
  addBox(size, options) {
    const o = options || {};
    const s = vector(size, 3, 'size');
    s.forEach((v) => positive(v, 'size'));
    const density = positive(o.density === undefined ? 1 : o.density, 'density', true);
    const friction = positive(o.friction === undefined ? 0.6 : o.friction, 'friction', true);
    const pos = vector(o.position || [0, 0, 0], 3, 'position');
    const [m, moment, radius] = boxMass(s, density);
    return this.add(bodyRecord(pos, s, m, moment, radius, friction, SHAPE_BOX), o);
  }

  /** A solid sphere (scenes.js:145-159; shapes.ts:28-36). */
  addSphere(radius, options) {
    const o = options || {};
    const r = positive(radius, 'radius');
    const density = positive(o.density === undefined ? 1 : o.density, 'density', true);
    const friction = positive(o.friction === undefined ? 0.6 : o.friction, 'friction', true);
    const pos = vector(o.position || [0, 0, 0], 3, 'position');
    const m = density > 0 ? (4 / 3) * Math.PI * r * r * r * density : 0;
    const inertia = 0.4 * m * r * r;
    return this.add(bodyRecord(pos, [2 * r, 2 * r, 2 * r], m, [inertia, inertia, inertia], r, friction, SHAPE_SPHERE), o);
  }

  /** An analytic capsule along local Y; height is the full tip-to-tip extent (scenes.js:181-197). */
  addCapsule(radius, height, options) {
    const o = options || {};
    const r = Number(radius);
    const h = Number(height);
    if (!(Number.isFinite(r) && r > 0 && Number.isFinite(h) && h >= 2 * r))
      throw Error('Capsule height must be at least twice its positive radius');
    const density = positive(o.density === undefined ? 1 : o.density, 'density', true);
    const friction = positive(o.friction === undefined ? 0.6 : o.friction, 'friction', true);
    const pos = vector(o.position || [0, 0, 0], 3, 'position');
    const [m, moment, bound] = capsuleMass(r, h, density);
    return this.add(bodyRecord(pos, [2 * r, h, 2 * r], m, moment, bound, friction, SHAPE_CAPSULE), o);
  }

  /** A convex hull of xyz points, its principal frame at `position` (its centre of mass) (scenes.js:160-179). */
  addHull(points, options) {
    this.assertEditable();
    const shape = convexHull(flatPoints(points, 'Hull points'));
    if (!shape) throw Error('A convex hull requires finite points spanning a volume');
    return this.addHullShape(shape, options);
  }

  /** A body of a cooked HullShape (shapes.ts:50-70 `hull`). */
  addHullShape(shape, options) {
    const o = options || {};
    this.assertEditable();
    if (shape.vertices.length / 3 > MAX_HULL_VERTICES || shape.faces.length > 255 || shape.edges.length > 255)
      throw Error(`hull: ${shape.vertices.length / 3} vertices, ${shape.faces.length} faces, ${shape.edges.length} edges (at most ${MAX_HULL_VERTICES} vertices, 255 faces and edges)`);
    const density = positive(o.density === undefined ? 1 : o.density, 'density', true);
    const friction = positive(o.friction === undefined ? 0.6 : o.friction, 'friction', true);
    const pos = vector(o.position || [0, 0, 0], 3, 'position');
    if (!this.hulls.has(shape)) {
      this.hulls.set(shape, this.hullWords);
      this.hullWords += shape.words;
    }
    const at = this.hulls.get(shape);
    const m = density > 0 ? shape.volume * density : 0;
    const moment = shape.moments.map((x) => x * Math.max(density, 0));
    if (Math.fround(SHAPE_HULL + at) !== SHAPE_HULL + at) throw Error('hull buffer beyond 2^24 words');
    return this.add(bodyRecord(pos, shape.size, m, moment, shape.radius, friction, SHAPE_HULL + at), o);
  }

  /** The scene's hull records, concatenated (4 u32 per word), or null without hulls. */
  hullBuffer() {
    if (!this.hulls.size) return null;
    const out = new Uint32Array(4 * this.hullWords);
    for (const [shape, at] of this.hulls) out.set(shape.record(at), 4 * at);
    return out;
  }

  addHeightfield(options) { return this.addTerrain(cookHeightfield(options), options); }

  addStaticMesh(options) { return this.addTerrain(cookStaticMesh(options), options); }

  /** Register cooked terrain pieces as static hull bodies (scenes.js:203-231). */
  addTerrain(cooked, options) {
    this.assertEditable();
    const bodies = createTerrainBodies(this, cooked, options);
    return { bodies, triangleCount: cooked.triangleCount, kind: cooked.kind };
  }

  check(a, b) {
    const n = this.records.length;
    if (b === null || b === undefined || a === b || !(b >= 0 && b < n) || (a !== null && !(a >= 0 && a < n)))
      throw Error('Constraint endpoints must be distinct bodies in this scene');
  }

  /** Ball joint plus optional angle lock (scenes.js:508-540; forces.ts:68-110). a = null: anchorA is a world point. */
  addJoint(a, b, options) {
    const o = options || {};
    this.assertEditable();
    this.check(a, b);
    const linear = o.linearStiffness === undefined ? Infinity : o.linearStiffness;
    const angular = o.angularStiffness === undefined ? 0 : o.angularStiffness;
    const breakForce = o.breakForce === undefined ? Infinity : o.breakForce;
    for (const v of [linear, angular, breakForce]) if (v !== Infinity) positive(v, 'joint stiffness/breakForce', true);
    const rA = vector(o.anchorA || [0, 0, 0], 3, 'anchorA');
    const rB = vector(o.anchorB || [0, 0, 0], 3, 'anchorB');
    const frac = o.breakOnPull && breakForce !== Infinity ? -breakForce : breakForce;
    const sA = a !== null ? this.records[a].subarray(16, 19) : [0, 0, 0];
    const sB = this.records[b].subarray(16, 19);
    const s = [sA[0] + sB[0], sA[1] + sB[1], sA[2] + sB[2]];
    const rec = new Float32Array(32);
    rec[3] = finite(linear);
    rec[7] = finite(angular);
    rec[11] = finite(frac);
    rec[15] = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
    rec.set(rA, 24);
    rec.set(rB, 28);
    this.joints.push([rec, [T_JOINT, a === null ? -1 : a, b, 0]]);
    return this.joints.length - 1;
  }

  /** Spring between body-local anchors; rest -1 takes their initial separation (scenes.js:546-575). */
  addSpring(a, b, options) {
    const o = options || {};
    this.assertEditable();
    this.check(a, b);
    if (a === null) throw Error('A spring requires two bodies');
    const stiffness = positive(o.stiffness === undefined ? 1000 : o.stiffness, 'stiffness', true);
    let rest = o.rest === undefined ? -1 : o.rest;
    if (rest !== -1) positive(rest, 'rest', true);
    const rA = vector(o.anchorA || [0, 0, 0], 3, 'anchorA');
    const rB = vector(o.anchorB || [0, 0, 0], 3, 'anchorB');
    if (rest < 0) {
      const ra = this.records[a], rb = this.records[b];
      const qa = Array.from(ra.subarray(4, 8)), qb = Array.from(rb.subarray(4, 8));
      const pa = rotate(qa, rA).map((x, k) => x + ra[k]);
      const pb = rotate(qb, rB).map((x, k) => x + rb[k]);
      const d = [pa[0] - pb[0], pa[1] - pb[1], pa[2] - pb[2]];
      rest = Math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    }
    const rec = new Float32Array(32);
    rec[3] = stiffness;
    rec.set(rA, 24);
    rec[27] = rest;
    rec.set(rB, 28);
    this.joints.push([rec, [T_SPRING, a, b, 0]]);
    return this.joints.length - 1;
  }

  /** The pair never collides (ref/forces.ts IgnoreCollision). */
  ignoreCollision(a, b) {
    this.assertEditable();
    this.check(a, b);
    this.ignore.push([a, b]);
  }

  /** Packed records in the layouts of include/avbd.h (what avbd_world_create takes). */
  packed() {
    const n = this.records.length;
    const bodies = new Float32Array(40 * n);
    this.records.forEach((r, i) => bodies.set(r, 40 * i));
    const joints = new Float32Array(32 * this.joints.length);
    const jointInfo = new Int32Array(4 * this.joints.length);
    this.joints.forEach(([r, info], j) => { joints.set(r, 32 * j); jointInfo.set(info, 4 * j); });
    const ignore = new Uint32Array(2 * this.ignore.length);
    this.ignore.forEach(([a, b], k) => { ignore[2 * k] = a; ignore[2 * k + 1] = b; });
    let filters = null;
    if (this.filters.size) {
      filters = new Uint32Array(2 * n).fill(0xffffffff);
      for (const [i, [g, m]] of this.filters) { filters[2 * i] = g; filters[2 * i + 1] = m; }
    }
    return Object.assign({ bodies, joints, jointInfo, ignore, filters, hulls: this.hullBuffer(), up: this.up }, this.params);
  }
}

function nativeOptions(options) {
  const o = options || {};
  const bp = o.broadphase === undefined ? 'auto' : o.broadphase;
  const code = { auto: 0, grid: 1, hploc: 2 }[bp];
  if (code === undefined) throw Error(`Unknown broadphase: ${bp}`);
  const capacity = o.capacity || {};
  return {
    colorCap: o.colorCap || 0,
    colorRounds: o.colorRounds || 0,
    pairCapacity: capacity.pairs || 0,
    graph: o.graph ? 1 : 0,
    adaptive: o.adaptive ? 1 : 0,
    broadphase: code,
    bvhRebuildInterval: (o.bvh && o.bvh.rebuildInterval) || 0,
    solver: solverCode(o.solver),
    worldBlock: o.worldBlock || 0,
    dataflowShare: o.dataflowShare || 0,
  };
}

// --- body edits (src/native/runtime.js:16-175, solver3D.js:130-213, packedPoses.js:3-72) ---------------------------

const EDIT_KEYS = ['linearVelocity', 'angularVelocity', 'impulse', 'force', '', 'angularImpulse', 'torque'];
const TELEPORT = 4;

/** An {x, y, z[, w]} object or an array-like of exactly `count` finite numbers (runtime.js:16-32). */
function components(value, count, name) {
  const a = value && typeof value.x === 'number'
    ? [value.x, value.y].concat(count >= 3 ? [value.z] : [], count === 4 ? [value.w] : [])
    : Array.from(value || []);
  if (a.length !== count || !a.every(Number.isFinite)) throw Error(`${name} requires ${count} finite components`);
  return a;
}

function unitRotation(rotation) {
  const q = components(rotation, 4, 'rotation');
  const l = Math.hypot(q[0], q[1], q[2], q[3]);
  if (!l) throw Error('rotation must be nonzero');
  return q.map((v) => v / l);
}

function coefficient(v) {
  if (typeof v !== 'number' || !Number.isFinite(v) || v < 0 || v > 1) throw Error('restitution must be between 0 and 1');
  return v;
}

function mask32(v) {
  if (!Number.isInteger(v) || v < 0 || v > 0xffffffff) throw Error('Collision masks must be unsigned 32-bit integers');
  return v;
}

/** Growable 12-word command records: body, kind, has point, 0 | value xyzw (f32) | point xyzw (f32). */
class CommandQueue {
  constructor() { this.count = 0; this.u32 = new Uint32Array(64 * 12); this.f32 = new Float32Array(this.u32.buffer); }
  reserve(n) {
    if (n * 12 <= this.u32.length) return;
    const u = new Uint32Array(Math.max(n, 2 * this.u32.length / 12) * 12);
    u.set(this.u32.subarray(0, this.count * 12));
    this.u32 = u;
    this.f32 = new Float32Array(u.buffer);
  }
  append(index, kind, value, point, hasPoint) {
    this.reserve(this.count + 1);
    const o = this.count++ * 12;
    this.u32.fill(0, o, o + 12);
    this.u32[o] = index;
    this.u32[o + 1] = kind;
    this.u32[o + 2] = hasPoint ? 1 : 0;
    for (let k = 0; k < 4; k++) {
      this.f32[o + 4 + k] = value[k] === undefined ? 0 : value[k];
      this.f32[o + 8 + k] = point && point[k] !== undefined ? point[k] : 0;
    }
  }
  extend(other) {
    this.reserve(this.count + other.count);
    this.u32.set(other.u32.subarray(0, other.count * 12), this.count * 12);
    this.count += other.count;
  }
  records() { return this.u32.subarray(0, this.count * 12); }
}

function solverCode(solver) {
  const code = { auto: 0, colors: 1, dataflow: 2, worlds: 3 }[solver === undefined ? 'auto' : solver];
  if (code === undefined) throw Error(`Unknown solver: ${solver}`);
  return code;
}

/** A world (or a batch of worlds) on the GPU with the shape of AvbdGpuSolver3D (scenes.d.ts:317-454). */
/**
 * A body handle (runtime.js:168-237 NativeBody): chainable edits of one body slot, and readState(). Triggers are not
 * built on this backend: turning one on throws, turning it off is accepted.
 */
class HipBody {
  constructor(solver, index) {
    this.solver = solver;
    this.index = index;
  }
  get gpuIndex() { return this.solver.liveIndex(this.index); }
  setLinearVelocity(v) { this.solver.setLinearVelocity(this.gpuIndex, v); return this; }
  setAngularVelocity(v) { this.solver.setAngularVelocity(this.gpuIndex, v); return this; }
  applyImpulse(v, p) { this.solver.applyImpulse(this.gpuIndex, v, p); return this; }
  applyForce(v, p) { this.solver.applyForce(this.gpuIndex, v, p); return this; }
  applyAngularImpulse(v) { this.solver.applyAngularImpulse(this.gpuIndex, v); return this; }
  applyTorque(v) { this.solver.applyTorque(this.gpuIndex, v); return this; }
  teleport(p, r) { this.solver.teleport(this.gpuIndex, p, r); return this; }
  setTrigger(enabled) {
    this.gpuIndex;
    if (enabled) throw Error('Trigger bodies are not available on the hip backend');
    return this;
  }
  setRestitution(value) { this.solver.setRestitution(this.gpuIndex, value); return this; }
  setCollisionGroups(group, collidesWith) {
    this.solver.setFilters([this.gpuIndex], [group], [collidesWith === undefined ? 0xffffffff : collidesWith]);
    return this;
  }
  setSleepEnabled(enabled) { this.solver.setSleepEnabled(this.gpuIndex, enabled); return this; }
  wakeUp() { this.solver.wakeBodies([this.gpuIndex]); return this; }
  readState() { return this.solver.readBodyState(this.gpuIndex); }
}

class HipSolver3D {
  constructor(scenes, options, device) {
    const list = Array.isArray(scenes) ? scenes : [scenes];
    if (!list.length) throw Error('An AVBD scene builder is required');
    for (const s of list) {
      if (!(s instanceof AvbdScene3D)) throw Error('An AVBD scene builder is required');
      s.assertEditable();
    }
    const opts = Object.assign(nativeOptions(options), { device });
    this.handle = native().createWorld(list.map((s) => s.packed()), opts);
    for (const s of list) s.created = true;
    const counts = native().counts(this.handle);
    this.bodyCount = counts.bodies;
    this.jointCount = counts.joints;
    this.params = Object.assign({}, list[0].params, { up: list[0].up.slice() });
    this.timeAccumulator = 0;
    this.destroyed = false;
    this.firstBodies = [];
    this.dts = [];
    let first = 0;
    for (const s of list) {
      this.firstBodies.push(first);
      this.dts.push(s.params.dt);
      first += s.bodyCount;
    }
    this.commands = new CommandQueue();
    this.handles = new Map();
    this.sceneBinding = null;
    this.scenes = list;
    // Bodies built with a restitution coefficient (scenes.js:684-687 applies them after creation the same way)
    const bouncy = [];
    list.forEach((sc, k) => {
      for (const [i, e] of Array.from(sc.restitution.entries()).sort((x, y) => x[0] - y[0]))
        bouncy.push([this.firstBodies[k] + i, e]);
    });
    if (bouncy.length) this.setRestitution(bouncy.map((b) => b[0]), bouncy.map((b) => b[1]));
    this.masses = new Float32Array(this.bodyCount);
    list.forEach((sc, k) => sc.records.forEach((r, i) => { this.masses[this.firstBodies[k] + i] = r[19]; }));
    // Sleeping (scenes.js:689-693): asked for in the options, or a body was built with allowSleep true
    const noSleep = [];
    list.forEach((sc, k) => { for (const i of Array.from(sc.noSleep).sort((x, y) => x - y)) noSleep.push(this.firstBodies[k] + i); });
    if (noSleep.length) native().setSleepEnabled(this.handle, Uint32Array.from(noSleep), new Uint8Array(noSleep.length));
    const sleeping = (options || {}).sleeping;
    if (sleeping || list.some((s) => s.sleepOptIn))
      this.enableSleeping(sleeping && typeof sleeping === 'object' ? sleeping : {});
  }

  // --- sleeping (worldSleeping.js; solver3D.js:317-343; appGpuSolver3D.js:869-889) ---------------------------------

  /** enableSleeping({speedThreshold, timeThreshold, islands, startAsleep}): bodies quiet and supported for
   * timeThreshold seconds sleep as fixed supports until an impact, an edit or a request wakes them; islands (default:
   * worlds of 4096 bodies or more) wakes and sleeps connected groups together; startAsleep (island mode) takes one
   * step that puts still, grounded groups to sleep at once. */
  enableSleeping(options) {
    this.alive();
    const o = options || {};
    for (const key of Object.keys(o))
      if (!['speedThreshold', 'timeThreshold', 'islands', 'startAsleep'].includes(key))
        throw Error(`Unknown sleep option: ${key}`);
    const speed = o.speedThreshold === undefined ? 0.03 : o.speedThreshold;
    const delay = o.timeThreshold === undefined ? 0.5 : o.timeThreshold;
    if (!Number.isFinite(speed) || speed <= 0 || !Number.isFinite(delay) || delay <= 0)
      throw Error('Sleep speedThreshold and timeThreshold must be positive and finite');
    if (o.startAsleep !== undefined && typeof o.startAsleep !== 'boolean') throw Error('startAsleep must be a boolean');
    if (o.islands !== undefined && typeof o.islands !== 'boolean') throw Error('Sleep islands must be a boolean');
    if (this.sleeping) return this;
    this.flushEdits();
    native().enableSleeping(this.handle, speed, delay, o.islands === undefined ? -1 : Number(o.islands),
                            !!o.startAsleep);
    if (o.startAsleep && this.sleepIslands) this.step();
    return this;
  }
  get sleeping() { return !this.destroyed && native().sleepMode(this.handle).enabled; }
  get sleepIslands() { return !this.destroyed && native().sleepMode(this.handle).islands; }
  /** Every body wakes with its mass back (poses kept) and sleeping stops. */
  disableSleeping() { this.alive(); this.flushEdits(); native().disableSleeping(this.handle); return this; }
  wakeAll() { this.alive(); this.flushEdits(); native().wakeAll(this.handle); return this; }
  /** wakeBody for each index: held awake through the next step; in island mode their islands wake. */
  wakeBodies(indices) {
    this.alive();
    const idx = Uint32Array.from(indices, (i) => this.liveIndex(i));
    this.flushEdits();
    native().wakeBodies(this.handle, idx);
    return this;
  }
  /** Sleep eligibility of one body; allowing it enables sleeping; every body wakes. */
  setSleepEnabled(index, enabled) {
    this.alive();
    if (typeof enabled !== 'boolean') throw Error('Sleep flag must be boolean');
    const idx = Uint32Array.of(this.liveIndex(index));
    this.flushEdits();
    native().setSleepEnabled(this.handle, idx, Uint8Array.of(enabled ? 1 : 0));
    return this;
  }
  // --- shape queries (appGpuSolver3D.js:306-332; shapeQueries.js:99-258) --------------------------------------------

  async raycast(origin, direction, options) {
    const o = options || {};
    const [hit] = await this.raycastAll([{ origin, direction, maxDistance: o.maxDistance }], o);
    return hit;
  }
  /** Rays {origin, direction, maxDistance} against every body's current pose: per ray null or
   * {index, distance, normal, center, point}. Options: ignore, collidesWith, membership, includeTriggers, maxDistance. */
  raycastAll(rays, options) { return this.cast(rays.map((r) => Object.assign({}, r, { radius: 0 })), options); }
  async sphereCast(origin, radius, direction, options) {
    const o = options || {};
    const [hit] = await this.sphereCastAll([{ origin, radius, direction, maxDistance: o.maxDistance }], o);
    return hit;
  }
  sphereCastAll(casts, options) { return this.cast(casts, options); }
  cast(queries, options) {
    this.alive();
    const o = options || {};
    const ignore = o.ignore === undefined ? [] : Array.from(o.ignore);
    const mask = o.collidesWith === undefined ? 0xffffffff : o.collidesWith;
    if (!Number.isInteger(mask) || mask < 0 || mask > 0xffffffff) throw Error('Query mask must be an unsigned 32-bit integer');
    const include = o.includeTriggers === undefined ? true : o.includeTriggers;
    if (typeof include !== 'boolean') throw Error('includeTriggers must be boolean');
    if (o.membership !== undefined && (!Number.isInteger(o.membership) || o.membership < 0 || o.membership > 0xffffffff))
      throw Error('Query membership must be an unsigned 32-bit integer');
    if (ignore.length > 16 || ignore.some((i) => !Number.isInteger(i) || i < 0 || i >= this.bodyCount))
      throw Error('Queries can ignore up to 16 live body indices');
    if (queries.length > 65535) throw Error('Too many queries in one batch');
    const checked = queries.map((q) => {
      const origin = Array.from(q.origin), direction = Array.from(q.direction);
      const radius = q.radius === undefined ? 0 : q.radius;
      const maxDistance = q.maxDistance !== undefined ? q.maxDistance : o.maxDistance !== undefined ? o.maxDistance : 3e38;
      if (origin.length !== 3 || direction.length !== 3 || ![...origin, ...direction].every(Number.isFinite) ||
          Math.hypot(...direction) === 0)
        throw Error('A query requires 3 finite components and a nonzero direction');
      if (!Number.isFinite(radius) || radius < 0 || !Number.isFinite(maxDistance) || maxDistance < 0)
        throw Error('Query radius and maximum distance must be finite and nonnegative');
      const length = Math.hypot(...direction);
      return { origin, direction: direction.map((x) => x / length), radius, maxDistance };
    });
    if (!checked.length) return Promise.resolve([]);
    this.flushEdits();
    const data = new Float32Array(checked.length * 8);
    checked.forEach((q, i) => data.set([...q.origin, q.maxDistance, ...q.direction, q.radius], i * 8));
    const hits = native().cast(this.handle, data, Uint32Array.from(ignore), mask >>> 0,
                               o.membership === undefined ? -1 : o.membership);
    const index = new Uint32Array(hits.buffer);
    return Promise.resolve(checked.map((q, i) => {
      const distance = hits[i * 8 + 3];
      if (distance < 0) return null;
      const normal = Array.from(hits.subarray(i * 8, i * 8 + 3));
      const center = q.origin.map((x, k) => x + distance * q.direction[k]);
      return { index: index[i * 8 + 4], distance, normal, center, point: center.map((x, k) => x - q.radius * normal[k]) };
    }));
  }

  // --- contact events (solver3D.js:252-311; contacts.ts ContactWatch) -------------------------------------------

  /** Bounded contact-event storage, made on the first call (every body watched); indices: only these report. */
  watchContacts(options) {
    this.alive();
    const o = options || {};
    const maxPairs = o.maxPairs === undefined ? 8192 : o.maxPairs, maxEvents = o.maxEvents === undefined ? 4096 : o.maxEvents;
    for (const n of [maxPairs, maxEvents])
      if (!Number.isInteger(n) || n < 1) throw Error('Contact capacities must be positive integers');
    if (o.indices !== undefined && !Array.isArray(o.indices)) throw Error('Watched indices must be an array');
    const indices = o.indices === undefined ? undefined : o.indices.map((i) => this.liveIndex(i));
    this.flushEdits();
    native().watchContacts(this.handle, maxPairs, maxEvents);
    if (this.eventCapacity === undefined) this.eventCapacity = maxEvents;
    if (indices !== undefined) {
      const selected = new Set(indices);
      const all = Uint32Array.from({ length: this.bodyCount }, (_, i) => i);
      native().setWatched(this.handle, all, Uint8Array.from(all, (i) => (selected.has(i) ? 1 : 0)));
    }
    return this;
  }
  /** {events: [{type, a, b, step, point, impulse, normal, isTrigger}], dropped, full} since the last read. */
  readContactEvents() {
    this.alive();
    this.flushEdits();
    if (this.eventCapacity === undefined) this.watchContacts();
    const r = native().readContactEvents(this.handle, this.eventCapacity);
    const f = new Float32Array(r.words.buffer, r.words.byteOffset, r.words.length);
    const events = [];
    for (let e = 0; e < r.words.length / 12; e++) {
      const o = e * 12;
      events.push({ type: r.words[o] === 1 ? 'begin' : 'end', a: r.words[o + 1], b: r.words[o + 2], step: r.words[o + 3],
                    point: Array.from(f.subarray(o + 4, o + 7)), impulse: f[o + 7],
                    normal: Array.from(f.subarray(o + 8, o + 11)), isTrigger: false });
    }
    events.sort((x, y) => x.step - y.step);
    return Promise.resolve({ events, dropped: r.written - events.length, full: r.full });
  }
  /** New lifetimes for sorted unique slots: their pairs begin again and do not end. */
  resetContactHistory(indices) {
    this.alive();
    if (!(indices instanceof Uint32Array)) throw Error('Contact resets require Uint32 GPU slots');
    this.flushEdits();
    native().resetContactHistory(this.handle, indices);
    return this;
  }

  /** {sleeping, wakeRequested} of one world after the last step. */
  readSleepStats(world = 0) {
    this.alive();
    this.flushEdits();
    return Promise.resolve(native().sleepStats(this.handle, world));
  }

  alive() { if (this.destroyed) throw Error('The solver has been destroyed'); }

  /** The slot of a handle of this world, a scene builder's body index, or a slot (runtime.js:326-345). */
  bodyIndex(body) {
    this.alive();
    if (body instanceof HipBody) {
      if (body.solver !== this) throw Error('Body belongs to another solver');
      return body.gpuIndex;
    }
    return this.liveIndex(body);
  }
  /** The handle of a body (one per slot, created on first use). */
  body(body) {
    const index = this.bodyIndex(body);
    if (!this.handles.has(index)) this.handles.set(index, new HipBody(this, index));
    return this.handles.get(index);
  }
  gpuIndex(i) { return this.liveIndex(i); }

  /** readBodyState: position, rotation, linear and angular velocity of one body (after queued edits). */
  async readBodyState(index) {
    const i = this.liveIndex(index);
    const r = await this.readSelectedBodies([i]);
    const mass = this.masses[i];
    return { position: Array.from(r.subarray(0, 3)), rotation: Array.from(r.subarray(4, 8)),
             linearVelocity: Array.from(r.subarray(32, 35)), angularVelocity: Array.from(r.subarray(36, 39)),
             mass, effectiveMass: r[19], sleeping: mass > 0 && r[19] === 0 };
  }

  /** setVelocities(indices, {linear, angular}) (runtime.js:356-378): xyz triples per index, either or both. */
  setVelocities(indices, values) {
    this.alive();
    const v = values || {};
    if (v.linear === undefined && v.angular === undefined) throw Error('Provide linear and/or angular velocities');
    const list = Array.from(indices, (i) => this.liveIndex(i));
    // xyz per body: flat (3 numbers each) or one triple per body; the whole batch is checked before any is queued
    const take = (arr, name) => {
      if (arr === undefined) return undefined;
      const items = Array.from(arr);
      const flat = items.length && typeof items[0] === 'number' ? items : [].concat(...items.map((x) => Array.from(x)));
      if (flat.length !== list.length * 3 || !flat.every(Number.isFinite))
        throw Error(`${name} velocities must be finite xyz per body`);
      return flat;
    };
    const lin = take(v.linear, 'Linear'), ang = take(v.angular, 'Angular');
    const staged = new CommandQueue();
    list.forEach((index, k) => {
      if (lin) staged.append(index, 0, lin.slice(3 * k, 3 * k + 3));
      if (ang) staged.append(index, 1, ang.slice(3 * k, 3 * k + 3));
    });
    this.commands.extend(staged);
    return this;
  }
  setLinearVelocities(indices, values) { return this.setVelocities(indices, { linear: values }); }
  setAngularVelocities(indices, values) { return this.setVelocities(indices, { angular: values }); }


  /**
   * attachToScene(scene, {afterStep, maxSubSteps, maxFrameTime}) (runtime.js:275-311): step from the scene's
   * before-render observable with its engine's frame time; destroy with the scene.
   */
  attachToScene(scene, options) {
    this.alive();
    const o = Object.assign({}, options || {});
    const afterStep = o.afterStep;
    delete o.afterStep;
    if (!scene || !scene.onBeforeRenderObservable || !scene.onBeforeRenderObservable.add ||
        !scene.onDisposeObservable || !scene.onDisposeObservable.addOnce || !scene.getEngine)
      throw Error('A Babylon scene is required');
    if (afterStep !== undefined && typeof afterStep !== 'function') throw Error('afterStep must be a function');
    this.advance(0, o);  // validates the stepping limits before observers are registered
    if (this.sceneBinding) {
      if (this.sceneBinding.scene === scene) return this;
      throw Error('The solver is already attached to another scene');
    }
    const before = scene.onBeforeRenderObservable.add(() => {
      this.advance(Math.max(0, scene.getEngine().getDeltaTime() / 1000), o, afterStep);
    });
    const dispose = scene.onDisposeObservable.addOnce(() => this.destroy());
    this.sceneBinding = { scene, before, dispose };
    return this;
  }
  /** Remove the observers without destroying the world; reset the elapsed-time debt. */
  detachFromScene() {
    const b = this.sceneBinding;
    if (b) {
      b.scene.onBeforeRenderObservable.remove(b.before);
      b.scene.onDisposeObservable.remove(b.dispose);
      this.sceneBinding = null;
      this.timeAccumulator = 0;
    }
    return this;
  }

  /** Alive, with queued edits submitted (the reference flushes them before steps and reads). */
  ready() { this.alive(); this.flushEdits(); }

  /** Submit one step without waiting (solver3D.js:350-359); queued edits go first. */
  step() { this.ready(); native().step(this.handle, 1); }

  liveIndex(index) {
    this.alive();
    if (!Number.isInteger(index) || index < 0 || index >= this.bodyCount) throw Error('A live 3D GPU body index is required');
    return index;
  }

  dtOf(index) {
    let k = this.firstBodies.length - 1;
    while (k > 0 && this.firstBodies[k] > index) k--;
    return this.dts[k];
  }

  queueBodyCommand(index, kind, value, point, hasPoint) {
    this.commands.append(this.liveIndex(index), kind, value, point, hasPoint);
    return this;
  }
  setLinearVelocity(index, v) { return this.queueBodyCommand(index, 0, components(v, 3, 'velocity')); }
  setAngularVelocity(index, v) { return this.queueBodyCommand(index, 1, components(v, 3, 'angular velocity')); }
  applyImpulse(index, v, point) {
    const p = point === undefined ? undefined : components(point, 3, 'point');
    return this.queueBodyCommand(index, 2, components(v, 3, 'impulse'), p, p !== undefined);
  }
  applyForce(index, v, point) {
    const p = point === undefined ? undefined : components(point, 3, 'point');
    return this.queueBodyCommand(index, 3, components(v, 3, 'force').concat([this.dtOf(this.liveIndex(index))]), p,
      p !== undefined);
  }
  applyAngularImpulse(index, v) { return this.queueBodyCommand(index, 5, components(v, 3, 'angular impulse')); }
  applyTorque(index, v) {
    return this.queueBodyCommand(index, 6, components(v, 3, 'torque').concat([this.dtOf(this.liveIndex(index))]));
  }
  teleport(index, position, rotation) {
    const p = components(position, 3, 'position');
    const q = rotation === undefined ? undefined : unitRotation(rotation);
    return this.queueBodyCommand(index, TELEPORT, p, q, q !== undefined);
  }

  /** Packed teleports: unique ascending Uint32 slots, Float32 xyz positions, optional Float32 xyzw rotations. */
  setPoses(indices, positions, rotations) {
    this.alive();
    if (!(indices instanceof Uint32Array) || !(positions instanceof Float32Array) ||
        (rotations !== undefined && !(rotations instanceof Float32Array)))
      throw Error('Pose batches require Uint32 indices and Float32 pose arrays');
    if (positions.length !== indices.length * 3 || (rotations && rotations.length !== indices.length * 4))
      throw Error('Provide xyz positions and optional xyzw rotations per body');
    let previous = -1;
    for (let i = 0; i < indices.length; i++) {
      this.liveIndex(indices[i]);
      if (indices[i] <= previous) throw Error('Pose batch slots must be unique and ascending');
      previous = indices[i];
    }
    if (!positions.every(Number.isFinite)) throw Error('Pose positions must be finite');
    const inverse = [];
    for (let i = 0; rotations && i < indices.length; i++) {
      const l = Math.hypot(rotations[4 * i], rotations[4 * i + 1], rotations[4 * i + 2], rotations[4 * i + 3]);
      if (!Number.isFinite(l) || l <= 0) throw Error('Pose rotations must be finite and nonzero');
      inverse.push(1 / l);
    }
    const staged = new CommandQueue();
    for (let i = 0; i < indices.length; i++) {
      const p = [positions[3 * i], positions[3 * i + 1], positions[3 * i + 2]];
      const q = rotations ? [0, 1, 2, 3].map((k) => rotations[4 * i + k] * inverse[i]) : [0, 0, 0, 1];
      staged.append(indices[i], TELEPORT, p, q, !!rotations);
    }
    this.commands.extend(staged);
    return this;
  }

  /** editBodies (runtime.js:45-175): the batch is checked whole before anything is queued; per record a teleport
   * precedes velocity replacements, impulses and forces; group / collidesWith become setFilters. */
  editBodies(edits) {
    this.alive();
    const staged = new CommandQueue();
    const properties = [];
    for (const edit of edits) {
      const index = this.bodyIndex(edit.body);
      if (edit.isTrigger !== undefined) throw Error('isTrigger edits are not available on the hip backend');
      const restitution = edit.restitution === undefined ? undefined : coefficient(edit.restitution);
      if (edit.allowSleep !== undefined && typeof edit.allowSleep !== 'boolean')
        throw Error('allowSleep must be a boolean');
      if (edit.group !== undefined || edit.collidesWith !== undefined || restitution !== undefined ||
          edit.allowSleep !== undefined) {
        if (edit.collidesWith !== undefined && edit.group === undefined)
          throw Error('A collision-mask edit requires group as well');
        properties.push([index, edit.group === undefined ? undefined : mask32(edit.group),
          edit.collidesWith === undefined ? 0xffffffff : mask32(edit.collidesWith), restitution, edit.allowSleep]);
      }
      const point = edit.worldPoint === undefined ? undefined : components(edit.worldPoint, 3, 'worldPoint');
      if (edit.position !== undefined) {
        const q = edit.rotation === undefined ? undefined : unitRotation(edit.rotation);
        staged.append(index, TELEPORT, components(edit.position, 3, 'position'), q, q !== undefined);
      } else if (edit.rotation !== undefined) {
        throw Error('A rotation edit requires position; use teleport for a full pose');
      }
      for (let kind = 0; kind < EDIT_KEYS.length; kind++) {
        if (kind === TELEPORT || edit[EDIT_KEYS[kind]] === undefined) continue;
        const v = components(edit[EDIT_KEYS[kind]], 3, EDIT_KEYS[kind]);
        const withPoint = (kind === 2 || kind === 3) && point !== undefined;
        if (kind === 3 || kind === 6) v.push(this.dtOf(index));
        staged.append(index, kind, v, withPoint ? point : undefined, withPoint);
      }
    }
    this.commands.extend(staged);
    for (const [index, group, collides, restitution, allow] of properties) {
      if (group !== undefined) this.setFilters([index], [group], [collides]);
      if (restitution !== undefined) this.setRestitution(index, restitution);
      if (allow !== undefined) this.setSleepEnabled(index, allow);
    }
    return this;
  }

  /** setRestitution(index, value) (appGpuSolver3D.js:455-471), or many (indices, values): coefficients in [0, 1],
   * from the next step on; while any is above 0 each step ends with the bounce passes (src/gpu/restitution.js). */
  setRestitution(indices, values) {
    this.alive();
    const many = typeof indices !== 'number';
    const idx = Uint32Array.from(many ? indices : [indices], (i) => this.liveIndex(i));
    const val = Float32Array.from(many ? values : [values], coefficient);
    if (idx.length !== val.length) throw Error('Each body index needs a restitution value');
    this.flushEdits();  // commands queued before keep their place ahead of the change
    native().setRestitution(this.handle, idx, val);
    return this;
  }
  getRestitution() { this.alive(); return native().getRestitution(this.handle, this.bodyCount); }

  /** setFilters(indices, groups, collidesWith) (appGpuSolver3D.js:958-983). */
  setFilters(indices, groups, collidesWith) {
    this.alive();
    if (indices.length !== groups.length || indices.length !== collidesWith.length)
      throw Error('Each body index needs a group and collision mask');
    const idx = Uint32Array.from(indices, (i) => this.liveIndex(i));
    native().setFilters(this.handle, idx, Uint32Array.from(groups, mask32), Uint32Array.from(collidesWith, mask32));
    return this;
  }

  /** Submit queued body commands without stepping (solver3D.js:208-214). */
  flushEdits() {
    this.alive();
    if (this.commands.count) {
      const words = this.commands.records().slice();
      this.commands.count = 0;
      native().enqueueCommands(this.handle, words);
    }
    return this;
  }

  /** Step parameters of world k of the batch; setParams changes them from the next step on. */
  getParams(k) { this.alive(); return native().getParams(this.handle, k || 0); }
  setParams(changes, k) {
    this.ready();
    native().setParams(this.handle, k || 0, changes || {});
    const p = native().getParams(this.handle, k || 0);
    this.dts[k || 0] = p.dt;
    if (!k) Object.assign(this.params, { dt: p.dt, gravity: p.gravity, iterations: p.iterations });
    return this;
  }

  /** Fixed-step accumulator; returns the steps submitted (runtime.js:255-272). */
  advance(elapsed, options, afterStep) {
    this.alive();
    const o = options || {};
    const maxSubSteps = o.maxSubSteps === undefined ? 6 : o.maxSubSteps;
    const maxFrameTime = o.maxFrameTime === undefined ? 0.05 : o.maxFrameTime;
    if (!Number.isFinite(elapsed) || elapsed < 0) throw Error('elapsed must be finite nonnegative seconds');
    if (!Number.isInteger(maxSubSteps) || maxSubSteps < 1 || !Number.isFinite(maxFrameTime) || maxFrameTime <= 0)
      throw Error('maxSubSteps must be a positive integer and maxFrameTime positive');
    const dt = this.params.dt;
    this.timeAccumulator = Math.min(maxSubSteps * dt, this.timeAccumulator + Math.min(maxFrameTime, elapsed));
    let count = 0;
    while (this.timeAccumulator + 1e-12 >= dt && count < maxSubSteps) {
      this.step();
      this.timeAccumulator = Math.max(0, this.timeAccumulator - dt);
      count++;
      if (afterStep) afterStep(this);
    }
    return count;
  }

  /** Per-step device milliseconds of n steps (waits). */
  stepTimed(n) { this.ready(); return native().stepTimed(this.handle, n); }

  /** All bodies, 40 f32 each, in slot order (solver.ts:1421-1423). */
  readBodies() { this.ready(); return Promise.resolve(native().readBodies(this.handle, 0, this.bodyCount)); }

  /** readSelectedBodies (solver3D.js:223-230): 40 or (posesOnly) 8 f32 per body. */
  readSelectedBodies(indices, options) {
    this.ready();
    const idx = indices instanceof Uint32Array ? indices : Uint32Array.from(indices);
    return Promise.resolve(native().readSelected(this.handle, idx, !!(options && options.posesOnly)));
  }

  readJoints() { this.ready(); return Promise.resolve(native().readJoints(this.handle, 0, this.jointCount)); }

  /** {pairs, contacts, manifolds, joints, overflow, clashes, colors, sleeping} (world.js:672-692). */
  readCounters() {
    this.alive();
    const c = native().readCounters(this.handle);
    if (this.sleeping) c.sleeping = native().sleepStats(this.handle, 0).sleeping;
    return Promise.resolve(c);
  }

  /** The last step's broadphase pairs (a > b) as [a0, b0, a1, b1, ...] (solver.ts:1471-1475). */
  readPairs() { this.alive(); return Promise.resolve(native().readPairs(this.handle)); }

  /** readContactList (solver.ts:1435-1469): {a, b, feature, stick, rA, rB, pen, lam, normal} per contact point. */
  readContactList() {
    this.ready();
    const { manifolds, contacts } = native().readContacts(this.handle);
    const f32 = new Float32Array(contacts.buffer);
    const mf32 = new Float32Array(manifolds.buffer);
    const out = [];
    for (let m = 0; m < manifolds.length / 8; m++) {
      const a = manifolds[8 * m], b = manifolds[8 * m + 1], first = manifolds[8 * m + 2], count = manifolds[8 * m + 3] & 15;
      const normal = [mf32[8 * m + 4], mf32[8 * m + 5], mf32[8 * m + 6]];
      for (let c = first; c < first + count; c++) {
        const w = 16 * c;
        out.push({
          a, b, normal,
          feature: (contacts[w + 3] & 0x7fffffff) >>> 0,
          stick: (contacts[w + 3] & 0x80000000) !== 0,
          rA: [f32[w], f32[w + 1], f32[w + 2]],
          rB: [f32[w + 4], f32[w + 5], f32[w + 6]],
          pen: [f32[w + 8], f32[w + 9], f32[w + 10]],
          lam: [f32[w + 12], f32[w + 13], f32[w + 14]],
        });
      }
    }
    return Promise.resolve(out);
  }

  /** adapt(counters) (solver.ts:1484-1519): grow storage and the colour cap from the last counters. */
  adapt() { this.alive(); return native().adapt(this.handle); }

  destroy() {
    if (this.destroyed) return;
    this.detachFromScene();
    native().destroy(this.handle);
    this.destroyed = true;
    this.handles.clear();
  }
}

/**
 * createSolverPipeline({backend: "hip", device}) (pipeline.js:14-93): worlds share the pipeline's device; destroy()
 * releases its worlds. createBatch steps many scenes in one launch per phase.
 */
async function createSolverPipeline(options) {
  const o = options || {};
  const backend = o.backend === undefined ? 'hip' : o.backend;
  if (backend !== 'hip') throw Error(`Unsupported solver backend: ${backend}`);
  const device = o.device === undefined ? 0 : o.device;
  const worlds = new Set();
  let destroyed = false;
  const track = (world) => {
    Object.defineProperty(world, 'backend', { value: backend, configurable: true });
    worlds.add(world);
    const destroy = world.destroy;
    world.destroy = function () {
      destroy.call(this);
      worlds.delete(this);
    };
    return world;
  };
  return {
    backend,
    device,
    createSolver(scene, solverOptions) {
      if (destroyed) throw Error('The solver pipeline has been destroyed');
      if (!scene || typeof scene.assertEditable !== 'function') throw Error('An AVBD scene builder is required');
      return track(new HipSolver3D([scene], solverOptions, device));
    },
    createBatch(scenes, solverOptions) {
      if (destroyed) throw Error('The solver pipeline has been destroyed');
      return track(new HipSolver3D(scenes, solverOptions, device));
    },
    destroy() {
      if (destroyed) return;
      for (const world of Array.from(worlds)) world.destroy();
      worlds.clear();
      destroyed = true;
    },
  };
}

module.exports = { createSolverPipeline, AvbdScene3D, HipSolver3D, CommandQueue, native, FLAGS_DEFAULT, HullShape,
  MAX_HULL_VERTICES, convexHull, hullFromTriangles, cookHeightfield, cookStaticMesh, createTerrainBodies };

'use strict';
// Deterministic input harness and replay formats over the Node host.
//
//  * WaterDemo: substep-indexed water injection and cursor pushes, exact
//    snapshot/restore and .sunar record/replay — the timing semantics of the
//    reference's harness (src/demo_boot.js:71-419): an injection or push armed
//    for substep s is applied immediately BEFORE substep s advances; pushes in
//    arming order on the exact current state.
//  * .suna scene container and .sunar replay codec: the reference's JSON wire
//    formats (src/suna.js:46-124, src/sunar.js:27-92), base64 little-endian
//    word arrays, frame/order-sorted events.

const eng = require('./sunaengine');

const SUNA_MAGIC = 'SunaBox Scene';
const SUNAR_MAGIC = 'SunaBox Replay';
const DEMO_PARTICLE_CAP = 6000;
const SPEC_VERSION = 24;

const integer = (x, label, lo, hi) => {
  if (lo === undefined) lo = -2147483648;
  if (hi === undefined) hi = 2147483647;
  if (!Number.isInteger(x) || x < lo || x > hi) throw new Error(`${label} is outside ${lo}..${hi}`);
  return x;
};
const packWords = (a) => Buffer.from(a.buffer, a.byteOffset, a.byteLength).toString('base64');
function unpackWords(s, Type, words, label) {
  const b = Buffer.from(s || '', 'base64');
  if (b.byteLength !== words * 4) throw new Error(`${label} has ${b.byteLength} bytes; expected ${words * 4}`);
  const copy = new Uint8Array(b.byteLength);
  copy.set(b);
  return new Type(copy.buffer);
}

// ---- .suna (src/suna.js) ----------------------------------------------------
function encodeSuna(payload) {
  const { state, tints } = payload;
  if (!(state instanceof Int32Array)) throw new Error('Suna save needs an Int32Array state');
  if (!(tints instanceof Uint32Array)) throw new Error('Suna save needs a Uint32Array tint buffer');
  const n = integer(payload.scene && payload.scene.n, 'scene.n', 0, 131072);
  integer(payload.scene.nFluid, 'scene.nFluid', 0, n);
  if (state.length !== n * 8) throw new Error(`state has ${state.length} words; expected ${n * 8}`);
  if (tints.length !== n) throw new Error(`tints has ${tints.length} words; expected ${n}`);
  const doc = {
    magic: SUNA_MAGIC, formatVersion: 1, app: payload.app,
    scene: Object.assign({}, payload.scene, {
      state: { encoding: 'i32le/base64', words: state.length, data: packWords(state) },
      tints: { encoding: 'u32le/base64', words: tints.length, data: packWords(tints) },
    }),
    materials: payload.materials || { events: [] }, bodies: payload.bodies || [], drivers: payload.drivers || [],
    devices: payload.devices || [], settings: payload.settings || {},
  };
  return JSON.stringify(doc, null, 2) + '\n';
}
function decodeSuna(source) {
  const text = typeof source === 'string' ? source : Buffer.from(source).toString('utf8');
  let doc;
  try { doc = JSON.parse(text); } catch (e) { throw new Error(`not valid JSON: ${e.message}`); }
  if (!doc || doc.magic !== SUNA_MAGIC) throw new Error(`not a ${SUNA_MAGIC} file`);
  if (doc.formatVersion !== 1) throw new Error(`unsupported .suna formatVersion ${doc.formatVersion}; this build reads 1`);
  const s = doc.scene || {};
  const n = integer(s.n, 'scene.n', 0, 131072);
  integer(s.nFluid, 'scene.nFluid', 0, n);
  integer(s.domW, 'scene.domW', 1);
  integer(s.domH, 'scene.domH', 1);
  integer(s.cellShift, 'scene.cellShift', 1, 27);
  const state = unpackWords(s.state && s.state.data, Int32Array, n * 8, 'scene.state');
  const tints = unpackWords(s.tints && s.tints.data, Uint32Array, n, 'scene.tints');
  const scene = Object.assign({}, s);
  delete scene.state; delete scene.tints;
  return Object.assign({}, doc, { scene, state, tints });
}

// ---- .sunar (src/sunar.js) --------------------------------------------------
function checkEvents(events) {
  let prevFrame = -1, prevOrder = -1;
  events.forEach((e, i) => {
    e = e || {};
    const frame = integer(e.frame, `events[${i}].frame`, 0), order = integer(e.order, `events[${i}].order`, 0);
    if (frame < prevFrame || (frame === prevFrame && order <= prevOrder)) {
      throw new Error(`events[${i}] is not in stable frame/order sequence`);
    }
    if (typeof e.type !== 'string' || !e.type) throw new Error(`events[${i}].type is empty`);
    prevFrame = frame; prevOrder = order;
  });
}
function encodeSunar(replay) {
  if (!replay || !replay.initial) throw new Error('Suna replay needs an initial scene');
  const events = Array.isArray(replay.events) ? replay.events : [];
  checkEvents(events);
  const initial = JSON.parse(encodeSuna(replay.initial));
  const doc = {
    magic: SUNAR_MAGIC, formatVersion: 1, app: replay.app || initial.app,
    timing: {
      substepsPerFrame: integer(replay.substepsPerFrame !== undefined ? replay.substepsPerFrame : 2, 'substepsPerFrame', 1, 32),
      durationFrames: integer(replay.durationFrames || 0, 'durationFrames', 0),
      durationSubsteps: integer(replay.durationSubsteps || 0, 'durationSubsteps', 0),
    },
    initial, events, final: replay.final || null,
  };
  return JSON.stringify(doc, null, 2) + '\n';
}
function decodeSunar(source) {
  const text = typeof source === 'string' ? source : Buffer.from(source).toString('utf8');
  let doc;
  try { doc = JSON.parse(text); } catch (e) { throw new Error(`not valid JSON: ${e.message}`); }
  if (!doc || doc.magic !== SUNAR_MAGIC) throw new Error(`not a ${SUNAR_MAGIC} file`);
  if (doc.formatVersion !== 1) throw new Error(`unsupported .sunar formatVersion ${doc.formatVersion}; this build reads 1`);
  const t = doc.timing || {};
  integer(t.substepsPerFrame, 'timing.substepsPerFrame', 1, 32);
  integer(t.durationFrames, 'timing.durationFrames', 0);
  integer(t.durationSubsteps, 'timing.durationSubsteps', 0);
  const events = Array.isArray(doc.events) ? doc.events : [];
  checkEvents(events);
  return { app: doc.app, substepsPerFrame: t.substepsPerFrame, durationFrames: t.durationFrames,
    durationSubsteps: t.durationSubsteps, initial: decodeSuna(JSON.stringify(doc.initial)), events,
    final: doc.final || null };
}

// ---- WaterDemo (src/demo_boot.js:127-420) ----------------------------------
const wu = (x) => Math.round(x * eng.ONE) | 0;
function makeDroplet(cx, cy, w, h, vx, vy) {
  if (w === undefined) w = 3; if (h === undefined) h = 3;
  const out = [];
  const x0 = cx - ((w - 1) * eng.ONE >> 1), y0 = cy - ((h - 1) * eng.ONE >> 1);
  for (let gy = 0; gy < h; gy++)
    for (let gx = 0; gx < w; gx++) out.push({ x: (x0 + gx * eng.ONE) | 0, y: (y0 + gy * eng.ONE) | 0, vx: vx | 0, vy: vy | 0 });
  return out;
}
const int = (x, label) => { if (!Number.isInteger(x)) throw new Error(`[demo_boot] ${label} must be an integer, got ${x}`); return x; };

class WaterDemo {
  constructor(engine, cap) {
    this.engine = engine; this.cap = cap;
    this.substep = 0; this._schedule = new Map(); this._pushes = new Map(); this.pushHits = 0; this._rec = null;
  }
  get n() { return this.engine.n; }
  get nFluid() { return this.engine.nFluid; }
  async digest() { return (await this.engine.digest()).hex; }
  async chainHex() { return this.engine.chainHex(); }
  destroy() { this.engine.destroy(); }
  injectWater(ev) {
    const list = Array.isArray(ev) ? ev : [ev];
    for (const e of list) {
      const s = int(e.substep, 'event.substep');
      if (s < this.substep) throw new Error(`[demo_boot] injection at substep ${s} is in the past (now ${this.substep})`);
      const ps = (e.particles || []).map((p) => ({ x: int(p.x, 'particle.x'), y: int(p.y, 'particle.y'), vx: (p.vx || 0) | 0, vy: (p.vy || 0) | 0 }));
      if (!ps.length) continue;
      const b = this._schedule.get(s);
      if (b) b.push(...ps); else this._schedule.set(s, ps.slice());
      if (this._rec) this._rec.events.push({ frame: s - this._rec.startSubstep, order: this._rec.order++, type: 'inject', particles: ps });
    }
  }
  pushAt(ev) {
    const s = int(ev.substep, 'push.substep');
    if (s < this.substep) throw new Error(`[demo_boot] push at substep ${s} is in the past (now ${this.substep})`);
    const p = { x: int(ev.x, 'push.x'), y: int(ev.y, 'push.y'), r: int(ev.r, 'push.r'), ix: (ev.ix || 0) | 0, iy: (ev.iy || 0) | 0 };
    if (!p.ix && !p.iy) return;
    const b = this._pushes.get(s);
    if (b) b.push(p); else this._pushes.set(s, [p]);
    if (this._rec) this._rec.events.push(Object.assign({ frame: s - this._rec.startSubstep, order: this._rec.order++, type: 'push' }, p));
  }
  step(count) {
    if (count === undefined) count = 1;
    const end = this.substep + int(count, 'count');
    for (const s of this._pushes.keys()) {
      if (s >= this.substep && s < end) throw new Error(`[demo_boot] push armed at substep ${s}: use stepThrough()`);
    }
    return this._stepInjectOnly(count);
  }
  _stepInjectOnly(count) {
    let remaining = int(count, 'count');
    while (remaining > 0) {
      const due = this._schedule.get(this.substep);
      if (due) {
        this._schedule.delete(this.substep);
        this.engine.spawnFluid(due.map((p) => Object.assign({}, p, { mat: eng.MAT_WATER })), { noWeld: true });
      }
      let chunk = remaining;
      for (const s of this._schedule.keys()) if (s > this.substep && s - this.substep < chunk) chunk = s - this.substep;
      this.engine.step(chunk);
      this.substep += chunk; remaining -= chunk;
    }
    return this.substep;
  }
  async stepThrough(count) {
    if (count === undefined) count = 1;
    let remaining = int(count, 'count');
    while (remaining > 0) {
      const duePush = this._pushes.get(this.substep);
      if (duePush) {
        this._pushes.delete(this.substep);
        await this.engine.syncMirror();
        for (const p of duePush) this.pushHits += this.engine.applyImpulse(p.x, p.y, p.r, p.ix, p.iy);
      }
      let chunk = remaining;
      for (const s of this._pushes.keys()) if (s > this.substep && s - this.substep < chunk) chunk = s - this.substep;
      const target = this.substep + chunk;
      while (this.substep < target) {
        let inj = target - this.substep;
        for (const s of this._schedule.keys()) if (s > this.substep && s - this.substep < inj) inj = s - this.substep;
        this._stepInjectOnly(inj);
      }
      remaining -= chunk;
    }
    return this.substep;
  }
  async snapshot(opts) {
    const exact = !opts || opts.exact !== false;
    const e = this.engine;
    return {
      app: { name: 'SunaEngine demo', specVersion: SPEC_VERSION, demoParticleCap: this.cap },
      scene: { n: e.n, nFluid: e.nFluid, domW: e.domW, domH: e.domH, cellShift: e.cellShift, wallsOn: true },
      state: await e.readState(), tints: new Uint32Array(e.n),
      derived: exact ? await e.readDerived() : null, substep: this.substep,
    };
  }
  restore(snap) {
    let acc = null;
    if (snap.derived) {
      // the carried surface-tension acceleration (Derived words 8..9) is the
      // only scratch a substep reads from its predecessor
      const n = snap.scene.n;
      acc = new Int32Array(n * 2);
      for (let i = 0; i < n; i++) { acc[2 * i] = snap.derived[i * 16 + 8]; acc[2 * i + 1] = snap.derived[i * 16 + 9]; }
    }
    this._loadPayload(snap, acc);
    this.substep = snap.substep || 0;
  }
  _loadPayload(p, stAcc) {
    const s = p.scene || (p.initial && p.initial.scene);
    const state = p.state || (p.initial && p.initial.state);
    if (!s || !state) throw new Error('[demo_boot] payload has no scene/state');
    const n = s.n | 0;
    const su = new Uint32Array(state.buffer, state.byteOffset, state.length);
    const pos = new Int32Array(n * 2), vel = new Int32Array(n * 2), matId = new Uint32Array(n);
    const flags = new Uint32Array(n), temp = new Int32Array(n), fuel = new Int32Array(n);
    for (let i = 0; i < n; i++) {
      const o = i * 8;
      pos[2 * i] = state[o]; pos[2 * i + 1] = state[o + 1]; vel[2 * i] = state[o + 2]; vel[2 * i + 1] = state[o + 3];
      matId[i] = su[o + 4]; flags[i] = su[o + 5] & eng.FLAG_MASK; temp[i] = state[o + 6]; fuel[i] = state[o + 7];
    }
    this.engine.loadScene({ n, nFluid: s.nFluid, pos, vel, matId, flags, temp, fuel, domW: s.domW, domH: s.domH,
      cellShift: s.cellShift }, stAcc);
    this._schedule.clear(); this._pushes.clear(); this._rec = null; this.substep = 0;
  }
  async startRecording() {
    const startSubstep = this.substep;
    const initial = await this.snapshot({ exact: false });
    const pendingSchedule = new Map([...this._schedule.entries()].map(([s, ps]) => [s, ps.slice()]));
    const pendingPushes = new Map([...this._pushes.entries()].map(([s, ps]) => [s, ps.map((p) => Object.assign({}, p))]));
    this._loadPayload(initial);
    this._schedule = pendingSchedule; this._pushes = pendingPushes; this.substep = startSubstep;
    this._rec = { startSubstep, initial, events: [], order: 0 };
    for (const [s, ps] of [...this._schedule.entries()].sort((a, b) => a[0] - b[0])) {
      this._rec.events.push({ frame: s - startSubstep, order: this._rec.order++, type: 'inject', particles: ps.map((p) => Object.assign({}, p)) });
    }
    for (const [s, ps] of [...this._pushes.entries()].sort((a, b) => a[0] - b[0])) {
      for (const p of ps) this._rec.events.push(Object.assign({ frame: s - startSubstep, order: this._rec.order++, type: 'push' }, p));
    }
  }
  async stopRecording() {
    if (!this._rec) throw new Error('[demo_boot] stopRecording without startRecording');
    const rec = this._rec;
    this._rec = null;
    rec.events.sort((a, b) => (a.frame - b.frame) || (a.order - b.order));
    rec.events.forEach((e, i) => { e.order = i; });
    const durationSubsteps = this.substep - rec.startSubstep;
    const doc = { app: rec.initial.app, initial: rec.initial, events: rec.events, substepsPerFrame: 1,
      durationFrames: durationSubsteps, durationSubsteps,
      final: { digest: await this.digest(), n: this.n, nFluid: this.nFluid } };
    return { text: encodeSunar(doc), doc };
  }
}

async function bootWaterDemo(opts) {
  opts = opts || {};
  const maxParticles = opts.maxParticles || DEMO_PARTICLE_CAP;
  if (maxParticles > DEMO_PARTICLE_CAP) throw new Error(`[demo_boot] maxParticles ${maxParticles} exceeds the demo cap ${DEMO_PARTICLE_CAP}`);
  const scene = opts.scene || 'pool';
  const sc = (typeof scene === 'object') ? scene
    : scene === 'empty' ? eng.makeScene('empty_v1')
      : scene === 'pool' ? eng.makeScene('pool_v1', { w: opts.poolW || 60, h: opts.poolH || 20, mat: eng.MAT_WATER })
        : eng.makeScene(scene, opts.sceneOpts || {});
  if (sc.n > maxParticles) throw new Error(`[demo_boot] scene has ${sc.n} particles; the demo cap is ${maxParticles}`);
  for (let i = 0; i < sc.n; i++) if ((sc.matId[i] >>> 0) !== eng.MAT_WATER) throw new Error('[demo_boot] water-only harness: scene contains a non-WATER particle');
  const e = await eng.createEngine({ maxParticles });
  e.loadScene(sc);
  return new WaterDemo(e, maxParticles);
}

async function replayFromSunar(source) {
  const doc = decodeSunar(source);
  const cap = (doc.app && doc.app.demoParticleCap) || DEMO_PARTICLE_CAP;
  const e = await eng.createEngine({ maxParticles: cap });
  const demo = new WaterDemo(e, cap);
  demo._loadPayload(doc);
  for (const ev of doc.events) {
    if (ev.type === 'inject') demo.injectWater({ substep: ev.frame, particles: ev.particles });
    else if (ev.type === 'push') demo.pushAt({ substep: ev.frame, x: ev.x, y: ev.y, r: ev.r, ix: ev.ix, iy: ev.iy });
  }
  demo.replayDoc = doc;
  return demo;
}

module.exports = { WaterDemo, bootWaterDemo, replayFromSunar, makeDroplet, wu, encodeSuna, decodeSuna, encodeSunar, decodeSunar, DEMO_PARTICLE_CAP };

'use strict';
// The deterministic water harness over the Node host: the API of the
// reference's WaterDemo (src/demo_boot.js:71-420) with its input timing,
// snapshot/restore and .sunar record/replay, built on this host's own
// pieces.
//
// Timing contract (the reference's): an injection or push armed for substep
// s acts immediately before substep s advances; pushes at one substep act in
// arming order, before that substep's injection; a push selects fluid slots
// on the exact current state.
//
// How it runs here: every armed input sits in one timeline sorted by
// (substep, arming sequence).  Injections split a run into spans, because
// spawnFluid reloads the world.  Pushes inside a span do not: they are handed
// to the engine as a device-resident event schedule (Engine.setEvents) and
// applied inside the captured substeps, so a span of any length is one
// step() call.  A push due on the span's first substep is applied from the
// host first, which keeps "pushes before the injection" when both are due
// together.  pushHits is read from the engine's own count of moved particles.

const eng = require('./sunaengine');
const formats = require('./formats');

const DEMO_PARTICLE_CAP = 6000;
const SPEC_VERSION = 24;

const int = (x, label) => {
  if (!Number.isInteger(x)) throw new Error(`[demo_boot] ${label} must be an integer, got ${x}`);
  return x;
};

// World units to Q16.16 ticks, and a w x h droplet at 1 wu pitch centred
// on (cx, cy): the reference's input helpers (src/demo_boot.js:43-60).
const wu = (x) => Math.round(x * eng.ONE) | 0;
function makeDroplet(cx, cy, w, h, vx, vy) {
  int(cx, 'cx'); int(cy, 'cy');
  const W = w === undefined ? 3 : w, Hh = h === undefined ? 3 : h;
  const x0 = cx - ((W - 1) * eng.ONE >> 1), y0 = cy - ((Hh - 1) * eng.ONE >> 1);
  const out = [];
  for (let r = 0; r < Hh; r++) {
    for (let c = 0; c < W; c++) out.push({ x: (x0 + c * eng.ONE) | 0, y: (y0 + r * eng.ONE) | 0, vx: vx | 0, vy: vy | 0 });
  }
  return out;
}

// Armed inputs in (substep, seq) order; seq is the arming sequence.
class Timeline {
  constructor() { this.items = []; this.seq = 0; }
  arm(substep, kind, data) {
    const it = { substep, seq: this.seq++, kind, data };
    let lo = 0, hi = this.items.length;
    while (lo < hi) {
      const mid = (lo + hi) >> 1;
      if (this.items[mid].substep <= substep) lo = mid + 1; else hi = mid;
    }
    this.items.splice(lo, 0, it);
  }
  // Remove and return the `kind` items with substep in [lo, hi), in order.
  takeRange(lo, hi, kind) {
    const out = [], keep = [];
    for (const it of this.items) ((it.kind === kind && it.substep >= lo && it.substep < hi) ? out : keep).push(it);
    this.items = keep;
    return out;
  }
  // First substep after `now` holding an item of `kind` (Infinity if none).
  nextAfter(now, kind) {
    for (const it of this.items) if (it.substep > now && it.kind === kind) return it.substep;
    return Infinity;
  }
  any(lo, hi, kind) { return this.items.some((it) => it.kind === kind && it.substep >= lo && it.substep < hi); }
  clear() { this.items = []; }
}

class WaterDemo {
  constructor(engine, cap) {
    this.engine = engine;
    this.cap = cap;
    this.substep = 0;         // substep index since the last load or restore point
    this._base = 0;           // this.substep when the engine's own counter was 0
    this._inputs = new Timeline();
    this._hitsBefore = 0;     // pushHits of worlds loaded earlier in this demo
    this._rec = null;
  }
  get n() { return this.engine.n; }
  get nFluid() { return this.engine.nFluid; }
  // Particles moved by pushes over the demo's life (src/demo_boot.js:137).
  get pushHits() { return this._hitsBefore + this.engine.pushHits(); }
  async digest() { return (await this.engine.digest()).hex; }
  async chainHex() { return this.engine.chainHex(); }
  async debugCounters() { return Array.from(await this.engine.debugCounters()); }
  async checkErrors() { return this.engine.checkErrors(); }
  destroy() { this.engine.destroy(); }

  _journal(substep, type, fields) {
    if (!this._rec) return;
    this._rec.events.push(Object.assign({ frame: substep - this._rec.startSubstep, order: this._rec.order++, type },
      fields));
  }

  injectWater(events) {
    for (const ev of (Array.isArray(events) ? events : [events])) {
      const s = int(ev.substep, 'event.substep');
      if (s < this.substep) throw new Error(`[demo_boot] injection at substep ${s} is in the past (now ${this.substep})`);
      const particles = (ev.particles || []).map((p) => ({
        x: int(p.x, 'particle.x'), y: int(p.y, 'particle.y'), vx: (p.vx || 0) | 0, vy: (p.vy || 0) | 0,
      }));
      if (particles.length === 0) continue;
      this._inputs.arm(s, 'inject', particles);
      this._journal(s, 'inject', { particles });
    }
  }

  pushAt(ev) {
    const s = int(ev.substep, 'push.substep');
    if (s < this.substep) throw new Error(`[demo_boot] push at substep ${s} is in the past (now ${this.substep})`);
    const p = { x: int(ev.x, 'push.x'), y: int(ev.y, 'push.y'), r: int(ev.r, 'push.r'),
      ix: (ev.ix || 0) | 0, iy: (ev.iy || 0) | 0 };
    if (!p.ix && !p.iy) return;
    this._inputs.arm(s, 'push', p);
    this._journal(s, 'push', p);
  }

  // The reference's synchronous step refuses pushes (they need a synced
  // mirror there); the contract is kept.
  step(count) {
    const k = int(count === undefined ? 1 : count, 'count');
    const end = this.substep + k;
    for (const it of this._inputs.items) {
      if (it.kind === 'push' && it.substep >= this.substep && it.substep < end) {
        throw new Error(`[demo_boot] push armed at substep ${it.substep}: use stepThrough()`);
      }
    }
    return this._advance(k);
  }

  async stepThrough(count) { return this._advance(int(count === undefined ? 1 : count, 'count')); }

  _advance(count) {
    const end = this.substep + count;
    while (this.substep < end) {
      const now = this.substep;
      // pushes due now act before now's injection, from the host
      for (const it of this._inputs.takeRange(now, now + 1, 'push')) {
        const p = it.data;
        this.engine.applyImpulse(p.x, p.y, p.r, p.ix, p.iy);
      }
      const due = this._inputs.takeRange(now, now + 1, 'inject');
      if (due.length) {
        const list = [];
        for (const it of due) for (const p of it.data) list.push(Object.assign({}, p, { mat: eng.MAT_WATER }));
        this.engine.spawnFluid(list, { noWeld: true });   // refused past capacity, as the reference
      }
      const stop = Math.min(end, this._inputs.nextAfter(now, 'inject'));
      const pushes = this._inputs.takeRange(now + 1, stop, 'push');
      if (pushes.length) {
        this.engine.setEvents(pushes.map((it) => ({ substep: it.substep - this._base, cx: it.data.x, cy: it.data.y,
          r: it.data.r, ix: it.data.ix, iy: it.data.iy })));
      }
      this.engine.step(stop - now);
      if (pushes.length) this.engine.setEvents([]);
      this.substep = stop;
    }
    return this.substep;
  }

  // ---- keyframes (src/demo_boot.js:276-320) -------------------------------
  async snapshot(opts) {
    const exact = !opts || opts.exact !== false;
    const e = this.engine;
    return {
      app: { name: 'SunaEngine demo', specVersion: SPEC_VERSION, demoParticleCap: this.cap },
      scene: { n: e.n, nFluid: e.nFluid, domW: e.domW, domH: e.domH, cellShift: e.cellShift, wallsOn: true,
        substepCount: e.substepCount, formationClock: e.formationClock },
      state: await e.readState(),
      tints: await e.readTints(e.n),
      derived: exact ? await e.readDerived() : null,
      substep: this.substep,
    };
  }

  // With `derived` the continuation is exact; the carried surface-tension
  // acceleration (derived words 8..9) is the only scratch a substep reads
  // from its predecessor.  Clears armed inputs; the chain restarts.
  restore(snap) {
    let stAcc = null;
    if (snap.derived) {
      const n = snap.scene.n;
      stAcc = new Int32Array(n * 2);
      for (let i = 0; i < n; i++) {
        stAcc[2 * i] = snap.derived[i * 16 + 8];
        stAcc[2 * i + 1] = snap.derived[i * 16 + 9];
      }
    }
    this._loadPayload(snap, stAcc);
    this.substep = snap.substep || 0;
    this._base = this.substep;
  }

  _loadPayload(payload, stAcc) {
    const scene = payload.scene || (payload.initial && payload.initial.scene);
    const state = payload.state || (payload.initial && payload.initial.state);
    const tints = payload.tints || (payload.initial && payload.initial.tints);
    if (!scene || !state) throw new Error('[demo_boot] payload has no scene/state');
    const n = scene.n | 0;
    const u = new Uint32Array(state.buffer, state.byteOffset, state.length);
    const sc = { n, nFluid: scene.nFluid, domW: scene.domW, domH: scene.domH, cellShift: scene.cellShift,
      pos: new Int32Array(n * 2), vel: new Int32Array(n * 2), matId: new Uint32Array(n), flags: new Uint32Array(n),
      temp: new Int32Array(n), fuel: new Int32Array(n) };
    for (let i = 0; i < n; i++) {
      const w = i * 8;
      sc.pos[2 * i] = state[w]; sc.pos[2 * i + 1] = state[w + 1];
      sc.vel[2 * i] = state[w + 2]; sc.vel[2 * i + 1] = state[w + 3];
      sc.matId[i] = u[w + 4]; sc.flags[i] = u[w + 5] & eng.FLAG_MASK;
      sc.temp[i] = state[w + 6]; sc.fuel[i] = state[w + 7];
    }
    if (this.engine.n > 0) this._hitsBefore += this.engine.pushHits();
    this.engine.loadScene(sc, stAcc);
    if (tints) this.engine.setTints(0, tints);
    this._inputs.clear();
    this._rec = null;
    this.substep = 0;
    this._base = 0;
  }

  // ---- .sunar record / replay (src/demo_boot.js:322-419) -----------------
  // Recording starts on a load boundary: the live world is reloaded from its
  // own snapshot without scratch, exactly as a replay will start, and every
  // input still armed for the future is journalled first.
  async startRecording() {
    const startSubstep = this.substep;
    const initial = await this.snapshot({ exact: false });
    const pending = this._inputs.items.slice();
    const seq = this._inputs.seq;
    this._loadPayload(initial);
    this._inputs.items = pending;
    this._inputs.seq = seq;
    this.substep = startSubstep;
    this._base = startSubstep;
    this._rec = { startSubstep, initial, events: [], order: 0 };
    // pending injections, one journal event per substep (their particles in
    // arming order), then pending pushes, each in substep order
    const byStep = new Map();
    for (const it of pending) if (it.kind === 'inject') {
      if (!byStep.has(it.substep)) byStep.set(it.substep, []);
      byStep.get(it.substep).push(...it.data);
    }
    for (const s of Array.from(byStep.keys()).sort((a, b) => a - b)) {
      this._journal(s, 'inject', { particles: byStep.get(s).map((p) => Object.assign({}, p)) });
    }
    for (const it of pending) if (it.kind === 'push') this._journal(it.substep, 'push', Object.assign({}, it.data));
  }

  async stopRecording() {
    if (!this._rec) throw new Error('[demo_boot] stopRecording without startRecording');
    const rec = this._rec;
    this._rec = null;
    const events = rec.events.slice().sort((a, b) => (a.frame - b.frame) || (a.order - b.order));
    events.forEach((e, i) => { e.order = i; });
    const durationSubsteps = this.substep - rec.startSubstep;
    const doc = { app: rec.initial.app, initial: rec.initial, events, substepsPerFrame: 1,
      durationFrames: durationSubsteps, durationSubsteps,
      final: { digest: await this.digest(), n: this.n, nFluid: this.nFluid } };
    return { text: formats.encodeSunar(doc), doc };
  }
}

// bootWaterDemo (src/demo_boot.js:71-97): a water-only demo on a pool,
// empty or named scene, capped at DEMO_PARTICLE_CAP.
async function bootWaterDemo(opts) {
  const o = opts || {};
  const maxParticles = o.maxParticles || DEMO_PARTICLE_CAP;
  if (maxParticles > DEMO_PARTICLE_CAP) {
    throw new Error(`[demo_boot] maxParticles ${maxParticles} exceeds the demo cap ${DEMO_PARTICLE_CAP}`);
  }
  const which = o.scene || 'pool';
  let sc;
  if (typeof which === 'object') sc = which;
  else if (which === 'empty') sc = eng.makeScene('empty_v1');
  else if (which === 'pool') sc = eng.makeScene('pool_v1', { w: o.poolW || 60, h: o.poolH || 20, mat: eng.MAT_WATER });
  else sc = eng.makeScene(which, o.sceneOpts || {});
  if (sc.n > maxParticles) throw new Error(`[demo_boot] scene has ${sc.n} particles; the demo cap is ${maxParticles}`);
  for (let i = 0; i < sc.n; i++) {
    if ((sc.matId[i] >>> 0) !== eng.MAT_WATER) {
      throw new Error('[demo_boot] water-only harness: scene contains a non-WATER particle');
    }
  }
  const e = await eng.createEngine({ maxParticles });
  e.loadScene(sc);
  return new WaterDemo(e, maxParticles);
}

// replayFromSunar (src/demo_boot.js:104-125): the recording's initial world
// with its journal re-armed at substep = frame (substepsPerFrame 1).
async function replayFromSunar(source) {
  const doc = formats.decodeSunar(source);
  const cap = (doc.app && doc.app.demoParticleCap) || DEMO_PARTICLE_CAP;
  const demo = new WaterDemo(await eng.createEngine({ maxParticles: cap }), cap);
  demo._loadPayload(doc);
  for (const ev of doc.events) {
    if (ev.type === 'inject') demo.injectWater({ substep: ev.frame, particles: ev.particles });
    else if (ev.type === 'push') demo.pushAt({ substep: ev.frame, x: ev.x, y: ev.y, r: ev.r, ix: ev.ix, iy: ev.iy });
  }
  demo.replayDoc = doc;
  return demo;
}

module.exports = { WaterDemo, bootWaterDemo, replayFromSunar, makeDroplet, wu, DEMO_PARTICLE_CAP, Timeline };

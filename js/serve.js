'use strict';
// The serving loop of the JavaScript host: replay/verify jobs of mixed sizes
// served with continuous batching, over the N-API addon.
//
// A job is a seeded initial state (2D or 3D), a substep count and a stir
// schedule; its result is the per-substep H_set || H_slot stream.  Jobs
// arrive over time, wait in a queue, and are admitted into the smallest free
// partition of a resident world that holds them (admission by memory: each
// partition is a fixed slice of particle slots and grid cells).  A world
// advances all its running jobs by rounds of `chunk` substeps (fewer when
// every running job needs fewer); a job that ends inside a round runs on to
// the round's end, and its result is the first `substeps` entries of its
// stream, which the extra substeps do not touch.  A finished job's stream is
// read back and its partition retired.  One thread drives both worlds, one
// round on the device at a time, alternating between the worlds that have
// running jobs: each world's persistent kernels size their grids to the
// whole chip.  A scene's hashes depend only on its own slots, so a job's
// stream is the same alone and in any mix (tests/js/serve_gpu.js).

const sn = require('./sunaengine.js');
const { buildLuts3d } = require('./tables3d.js');

const ONE = sn.ONE;

// ---- seeded workloads -----------------------------------------------------
function splitmix32(a) {
  a = (a + 0x9E3779B9) >>> 0;
  let t = a;
  t = Math.imul(t ^ (t >>> 16), 0x21F0AAAD) >>> 0;
  t = Math.imul(t ^ (t >>> 15), 0x735A2D97) >>> 0;
  return (t ^ (t >>> 15)) >>> 0;
}

// A 0.5 wu water block of exactly nf particles, about 2:1 wide, standing on
// the floor of a walled 256 x 128 wu box (scaled up by powers of two until
// it fits), at x position xPick (half-wu steps, modulo the free width),
// jittered by integers in [-1/64, 1/64) wu from xorshift32(seed).
function fluidBlockScene(nf, seed, xPick) {
  const bw = Math.ceil(Math.sqrt(2 * nf)), bh = Math.ceil(nf / bw);
  let k = 1;
  while (256 * k < bw * 0.5 + 16 || 128 * k < bh * 0.5 + 24) k *= 2;
  const domW = 256 * k, domH = 128 * k;
  const maxX0 = Math.floor((domW - 6 - bw * 0.5) * 2);
  const x0 = 3 * ONE + ((xPick >>> 0) % Math.max(1, maxX0)) * (ONE >> 1);
  const rnd = sn.xorshift32(seed);
  const ys = (domH - 3) * ONE, pitch = ONE >> 1, jit = 1024;
  const bnd = sn.boundaryRing(domW, domH, domH - 12);
  const n = nf + bnd.length;
  const s = { n, nFluid: nf, pos: new Int32Array(2 * n), vel: new Int32Array(2 * n), matId: new Uint32Array(n),
              domW: domW * ONE, domH: domH * ONE, cellShift: 18 };
  let i = 0;
  for (let gy = 0; gy < bh; gy++) {
    for (let gx = 0; gx < bw; gx++) {
      const jx = (rnd() % (2 * jit)) - jit, jy = (rnd() % (2 * jit)) - jit;
      if (i < nf) {
        s.pos[2 * i] = x0 + gx * pitch + jx;
        s.pos[2 * i + 1] = ys - gy * pitch + jy;
        s.matId[i] = sn.MAT_WATER;
      }
      i++;
    }
  }
  bnd.forEach((b, j) => { s.pos[2 * (nf + j)] = b[0]; s.pos[2 * (nf + j) + 1] = b[1]; s.matId[nf + j] = sn.MAT_WATER; });
  return s;
}

// A side^3 block at `pitch` ticks from (1.5, 1.5, 1.5) wu in a g^3-cell world
// (cells of 2 wu), jittered per axis by integers in [-1/64, 1/64) wu from
// xorshift32(seed), x fastest, then z, then y; 16-word records
// {p.xyz, mat, v.xyz, 0, ...} with material 0 (water).
function blockScene3(side, g, seed) {
  const n = side * side * side, o = Math.round(1.5 * ONE), pitch = ONE >> 1, jit = 1024;
  const rnd = sn.xorshift32(seed);
  const parts = new Int32Array(n * 16);
  for (let c = 0; c < n; c++) {
    const jx = (rnd() % (2 * jit)) - jit, jy = (rnd() % (2 * jit)) - jit, jz = (rnd() % (2 * jit)) - jit;
    const i = c % side, kk = Math.floor(c / side) % side, j = Math.floor(c / (side * side));
    parts[16 * c] = o + i * pitch + jx;
    parts[16 * c + 1] = o + j * pitch + jy;
    parts[16 * c + 2] = o + kk * pitch + jz;
  }
  return { n, parts, gx: g, gy: g, gz: g };
}

// Stir impulses at Bernoulli(1/meanGap) substep indices: radius 8 wu,
// magnitude uniform in [0, 2] wu/substep in one of 16 (2D) or 14 (3D)
// directions given as integer unit vectors (x 2^16), so that no
// transcendental function enters a job's definition.
const DIR2 = [[65536, 0], [60547, 25080], [46341, 46341], [25080, 60547], [0, 65536], [-25080, 60547],
  [-46341, 46341], [-60547, 25080], [-65536, 0], [-60547, -25080], [-46341, -46341], [-25080, -60547],
  [0, -65536], [25080, -60547], [46341, -46341], [60547, -25080]];
const D3 = 37837;   // 2^16 / sqrt(3)
const DIR3 = [[65536, 0, 0], [-65536, 0, 0], [0, 65536, 0], [0, -65536, 0], [0, 0, 65536], [0, 0, -65536],
  [D3, D3, D3], [D3, D3, -D3], [D3, -D3, D3], [D3, -D3, -D3], [-D3, D3, D3], [-D3, D3, -D3], [-D3, -D3, D3],
  [-D3, -D3, -D3]];
const scaleDir = (mag, c) => Math.floor(mag * c / 65536);

function stirEvents2d(seed, substeps, box, meanGap) {
  meanGap = meanGap || 50;
  const rnd = sn.xorshift32(seed), [x0, y0, x1, y1] = box, out = [];
  for (let t = 0; t < substeps; t++) {
    if (rnd() % meanGap) continue;
    const cx = x0 + rnd() % Math.max(1, x1 - x0), cy = y0 + rnd() % Math.max(1, y1 - y0);
    const mag = rnd() % (2 * ONE + 1), d = DIR2[rnd() % DIR2.length];
    out.push([t, cx, cy, 8 * ONE, scaleDir(mag, d[0]), scaleDir(mag, d[1]), 3000]);
  }
  return out;
}

function stirEvents3d(seed, substeps, ext, meanGap) {
  meanGap = meanGap || 50;
  const rnd = sn.xorshift32(seed), out = [];
  for (let t = 0; t < substeps; t++) {
    if (rnd() % meanGap) continue;
    const c = [rnd() % Math.max(1, ext[0]), rnd() % Math.max(1, Math.floor(ext[1] / 2)), rnd() % Math.max(1, ext[2])];
    const mag = rnd() % (2 * ONE + 1), d = DIR3[rnd() % DIR3.length];
    out.push([t, c[0], c[1], c[2], 8 * ONE, scaleDir(mag, d[0]), scaleDir(mag, d[1]), scaleDir(mag, d[2]), 3000]);
  }
  return out;
}

function makeJob(jid, seed, kind, n, substeps, arrival) {
  let scene, events;
  if (kind === '2d') {
    scene = fluidBlockScene(n, seed, splitmix32(seed) & 0xFFFF);
    const box = [8 * ONE, Math.floor(scene.domH / 2), scene.domW - 8 * ONE, scene.domH - 8 * ONE];
    events = stirEvents2d((seed ^ 0x5A5A) >>> 0, substeps, box);
  } else {
    const side = Math.max(4, Math.round(Math.cbrt(n)));
    const g = Math.max(8, Math.floor((side + 1) / 4) + 6);
    scene = blockScene3(side, g, seed);
    events = stirEvents3d((seed ^ 0x5A5A) >>> 0, substeps, [g * 2 * ONE, g * 2 * ONE, g * 2 * ONE]);
  }
  return { jid, kind, scene, substeps, events, arrival: arrival || 0, n: scene.n,
           admitted: null, finished: null, slot: null, done: 0, stream: null };
}

// `count` jobs with exponential inter-arrival times at `rate` jobs/s: 70% 2D
// / 30% 3D, particle counts log-uniform in [nLo, nHi], substeps uniform in
// [stepsLo, stepsHi].
function poissonJobs(count, rate, seed, o) {
  o = Object.assign({ nLo: 4096, nHi: 262144, stepsLo: 100, stepsHi: 5000, frac3d: 0.3 }, o || {});
  const rnd = sn.xorshift32(seed);
  const jobs = [];
  let t = 0;
  for (let j = 0; j < count; j++) {
    const u = ((rnd() & 0xFFFFFF) + 0.5) / (1 << 24);
    t += -Math.log(u) / rate;
    const kind = (rnd() % 1000) < Math.floor(o.frac3d * 1000) ? '3d' : '2d';
    const v = (rnd() & 0xFFFFFF) / (1 << 24);
    const n = Math.round(Math.exp(Math.log(o.nLo) + v * (Math.log(o.nHi) - Math.log(o.nLo))));
    const steps = o.stepsLo + rnd() % (o.stepsHi - o.stepsLo + 1);
    jobs.push(makeJob(j, splitmix32((seed + 7919 * j) >>> 0), kind, n, steps, t));
  }
  return jobs;
}

// ---- worlds ---------------------------------------------------------------
const DEFAULT_2D = [[20000, 2048, 24], [70000, 8192, 16], [270000, 8192, 8]];
const DEFAULT_3D = [[20000, 8192, 12], [70000, 16384, 8], [270000, 65536, 4]];
let LUTS3 = null;

function footprint(job) {
  if (job.kind === '2d') {
    if (!job.image) job.image = sn.buildState(job.scene, 1 << 30, 1 << 30);
    return [job.image.n, job.image.gridW * job.image.gridH];
  }
  return [job.scene.n, job.scene.gx * job.scene.gy * job.scene.gz];
}

class World {
  constructor(kind, classes, o) {
    this.kind = kind;
    this.caps = [];
    for (const [cp, cc, count] of classes) for (let k = 0; k < count; k++) this.caps.push([cp, cc]);
    const maxP = this.caps.reduce((a, c) => a + c[0], 0), maxC = this.caps.reduce((a, c) => a + c[1], 0);
    const N = sn.native();
    this.N = N;
    const dbg = o.debug ? 1 : 0;
    if (kind === '2d') {
      this.h = N.create(maxP, maxC, this.caps.length, o.iters2d, o.streamCap, 0, sn.loadLuts(), sn.matsImage(), dbg);
      N.reserveSlots(this.h, Uint32Array.from(this.caps.map((c) => c[0])), Uint32Array.from(this.caps.map((c) => c[1])));
    } else {
      if (!LUTS3) LUTS3 = buildLuts3d();
      this.h = N.create3(maxP, maxC, this.caps.length, o.iters3d, o.streamCap, 0, LUTS3, dbg);
      N.reserveSlots3(this.h, Uint32Array.from(this.caps.map((c) => c[0])), Uint32Array.from(this.caps.map((c) => c[1])));
    }
    this.busy = this.caps.map(() => null);
    this.inflight = 0;
    this.rounds = 0;
  }
  destroy() {
    if (this.h) { this.kind === '2d' ? this.N.destroy(this.h) : this.N.destroy3(this.h); this.h = null; }
  }
  // the smallest free partition that holds the job
  findSlot(fp) {
    let best = -1;
    this.caps.forEach(([cp, cc], s) => {
      if (this.busy[s] === null && fp[0] <= cp && fp[1] <= cc && (best < 0 || cp < this.caps[best][0])) best = s;
    });
    return best;
  }
  fitsAnywhere(fp) { return this.caps.some(([cp, cc]) => fp[0] <= cp && fp[1] <= cc); }
  admit(slot, job) {
    if (this.kind === '2d') this.N.loadSlot(this.h, slot, job.image || sn.buildState(job.scene, 1 << 30, 1 << 30));
    else this.N.loadSlot3(this.h, slot, job.scene);
    this.busy[slot] = job;
  }
  setEvents() {
    const W = this.kind === '2d' ? 8 : 10, rows = [];
    this.busy.forEach((job, s) => { if (job) for (const e of job.events) rows.push([s].concat(e)); });
    rows.sort((a, b) => a[0] - b[0] || a[1] - b[1]);
    const arr = new Int32Array(Math.max(rows.length, 1) * W);
    rows.forEach((r, k) => arr.set(r, k * W));
    const words = arr.subarray(0, rows.length * W);
    if (this.kind === '2d') this.N.setEvents(this.h, words); else this.N.setEvents3(this.h, words);
  }
  setActive() {
    const m = Uint8Array.from(this.busy.map((j) => (j ? 1 : 0)));
    if (this.kind === '2d') this.N.setActive(this.h, m); else this.N.setActive3(this.h, m);
  }
  step(k) { if (this.kind === '2d') this.N.step(this.h, k); else this.N.step3(this.h, k); }
  query() { return this.kind === '2d' ? this.N.query(this.h) : this.N.query3(this.h); }
  sync() { if (this.kind === '2d') this.N.sync(this.h); else this.N.sync3(this.h); }
  stream(slot, substeps) {
    const s = this.kind === '2d' ? this.N.readHashStream(this.h, slot, substeps) : this.N.readHashStream3(this.h, slot, substeps);
    return s.slice(0, substeps * 8);
  }
}

// A short sleep of the host thread (the loop polls the device meanwhile).
const SLEEP = new Int32Array(new SharedArrayBuffer(4));
const sleepMs = (ms) => { if (ms > 0) Atomics.wait(SLEEP, 0, 0, ms); };

class Server {
  constructor(o) {
    o = Object.assign({ classes2d: DEFAULT_2D, classes3d: DEFAULT_3D, iters2d: 4, iters3d: 3, chunk: 64, overlap: true,
                        streamCap: 8192, debug: false }, o || {});
    this.o = o;
    this.worlds = {};
    if (o.classes2d.length) this.worlds['2d'] = new World('2d', o.classes2d, o);
    if (o.classes3d.length) this.worlds['3d'] = new World('3d', o.classes3d, o);
  }
  close() { for (const w of Object.values(this.worlds)) w.destroy(); this.worlds = {}; }

  // Serve `jobs` (arrival times in seconds from the start); with
  // realtime = false every job counts as already arrived.
  run(jobs, opts) {
    const realtime = !(opts && opts.realtime === false);
    const now = () => Number(process.hrtime.bigint()) / 1e9;
    for (const j of jobs) {
      const w = this.worlds[j.kind];
      if (!w || !w.fitsAnywhere(footprint(j))) throw new Error(`job ${j.jid}: no partition can hold it`);
      if (j.substeps + this.o.chunk > this.o.streamCap) throw new Error(`job ${j.jid}: substeps exceed the H ring`);
    }
    const kinds = Object.keys(this.worlds);
    const pending = {}, queue = {};
    for (const k of kinds) {
      pending[k] = jobs.filter((j) => j.kind === k).sort((a, b) => a.arrival - b.arrival);
      queue[k] = [];
    }
    const t0 = now();
    let turn = 0, current = null;
    const finishRound = (w) => {
      w.sync();
      const k = w.inflight;
      w.inflight = 0;
      w.rounds++;
      let retired = false;
      w.busy.forEach((j, s) => {
        if (!j) return;
        j.done += k;
        if (j.done >= j.substeps) {
          j.stream = w.stream(s, j.substeps);
          j.finished = now() - t0;
          w.busy[s] = null;
          retired = true;
        }
      });
      if (retired) w.setActive();
    };
    const admitAndLaunch = (kind) => {
      const w = this.worlds[kind];
      let admitted = false;
      for (const job of queue[kind].slice()) {
        const sl = w.findSlot(footprint(job));
        if (sl < 0) continue;
        w.admit(sl, job);
        job.slot = sl; job.admitted = now() - t0; job.done = 0;
        queue[kind].splice(queue[kind].indexOf(job), 1);
        admitted = true;
      }
      if (admitted) w.setEvents();
      const running = w.busy.filter((j) => j);
      if (running.length) {
        const k = Math.min(this.o.chunk, Math.max(...running.map((j) => j.substeps - j.done)));
        w.step(k);
        w.inflight = k;
      }
    };
    // overlap (the default): each world keeps a round in flight on its own
    // stream, so 2D and 3D rounds run side by side on the device
    while (this.o.overlap) {
      const t = now() - t0;
      for (const k of kinds) while (pending[k].length && (!realtime || pending[k][0].arrival <= t)) queue[k].push(pending[k].shift());
      for (const k of kinds) if (this.worlds[k].inflight && !this.worlds[k].query()) finishRound(this.worlds[k]);
      for (const k of kinds) if (!this.worlds[k].inflight) admitAndLaunch(k);
      if (kinds.some((k) => this.worlds[k].inflight)) { sleepMs(0.2); continue; }
      if (kinds.some((k) => queue[k].length)) throw new Error('queued jobs fit no free partition of an idle world');
      if (!kinds.some((k) => pending[k].length)) break;
      const next = Math.min(...kinds.filter((k) => pending[k].length).map((k) => pending[k][0].arrival));
      sleepMs(Math.min(5, (next - (now() - t0)) * 1e3));
    }
    for (;!this.o.overlap;) {
      const t = now() - t0;
      for (const k of kinds) while (pending[k].length && (!realtime || pending[k][0].arrival <= t)) queue[k].push(pending[k].shift());
      if (current !== null) {
        if (this.worlds[current].query()) { sleepMs(0.2); continue; }
        finishRound(this.worlds[current]);
        current = null;
      }
      for (let i = 0; i < kinds.length; i++) {
        const kind = kinds[(turn + i) % kinds.length];
        admitAndLaunch(kind);
        if (this.worlds[kind].inflight) { current = kind; turn = (kinds.indexOf(kind) + 1) % kinds.length; break; }
      }
      if (current === null) {
        if (kinds.some((k) => queue[k].length)) throw new Error('queued jobs fit no free partition of an idle world');
        if (!kinds.some((k) => pending[k].length)) break;
        const next = Math.min(...kinds.filter((k) => pending[k].length).map((k) => pending[k][0].arrival));
        sleepMs(Math.min(5, (next - (now() - t0)) * 1e3));
      }
    }
    return { wallS: now() - t0, rounds: Object.fromEntries(kinds.map((k) => [k, this.worlds[k].rounds])) };
  }
}

function percentile(xs, q) {
  const s = xs.slice().sort((a, b) => a - b);
  const pos = (s.length - 1) * q / 100, lo = Math.floor(pos), hi = Math.ceil(pos);
  return s[lo] + (s[hi] - s[lo]) * (pos - lo);
}

function summarize(jobs, wallS) {
  const lat = jobs.map((j) => j.finished - j.arrival), wait = jobs.map((j) => j.admitted - j.arrival);
  const work = jobs.reduce((a, j) => a + j.n * j.substeps, 0);
  const span = Math.max(...jobs.map((j) => j.finished)) - Math.min(...jobs.map((j) => j.arrival));
  return { jobs: jobs.length, particleSubsteps: work, throughput: span > 0 ? work / span : 0,
           p50LatencyS: percentile(lat, 50), p99LatencyS: percentile(lat, 99),
           p50WaitS: percentile(wait, 50), p99WaitS: percentile(wait, 99), wallS };
}

// One job alone in an engine of its own (the reference for the mix).
function runAlone(job, o) {
  o = Object.assign({ iters2d: 4, iters3d: 3, debug: true }, o || {});
  const N = sn.native(), dbg = o.debug ? 1 : 0;
  let out;
  if (job.kind === '2d') {
    const im = job.image || sn.buildState(job.scene, 1 << 30, 1 << 30);
    const h = N.create(im.n, im.gridW * im.gridH, 1, o.iters2d, job.substeps + 8, 0, sn.loadLuts(), sn.matsImage(), dbg);
    N.loadScenes(h, [im]);
    N.setEvents(h, Int32Array.from([].concat(...job.events.map((e) => [0].concat(e)))));
    N.step(h, job.substeps);
    out = N.readHashStream(h, 0, job.substeps);
    N.destroy(h);
  } else {
    if (!LUTS3) LUTS3 = buildLuts3d();
    const sc = job.scene;
    const h = N.create3(sc.n, sc.gx * sc.gy * sc.gz, 1, o.iters3d, job.substeps + 8, 0, LUTS3, dbg);
    N.load3(h, [sc]);
    N.setEvents3(h, Int32Array.from([].concat(...job.events.map((e) => [0].concat(e)))));
    N.step3(h, job.substeps);
    out = N.readHashStream3(h, 0, job.substeps);
    N.destroy3(h);
  }
  return out;
}

module.exports = { Server, World, makeJob, poissonJobs, summarize, runAlone, fluidBlockScene, blockScene3,
                   stirEvents2d, stirEvents3d, splitmix32, DEFAULT_2D, DEFAULT_3D };

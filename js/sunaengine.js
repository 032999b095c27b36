'use strict';
// Node.js host of the MI355X-native deterministic PBF engine.
//
// The reference engine's JavaScript API (src/engine.js: createEngine,
// Engine.loadScene / step / digest / chainHex / readState / readDerived /
// applyImpulse / spawnFluid / debugCounters / assertSpecValidRegime /
// syncMirror / destroy, and makeScene) over the thin N-API addon in
// addon.c, which calls the HIP library through its C ABI.  Plain ES2019 so
// that Node 12 runs it.  All compute happens in the HIP library.

const fs = require('fs');
const path = require('path');

const ROOT = path.resolve(__dirname, '..');
let addon = null;
function native() {
  if (!addon) {
    const p = path.join(__dirname, 'build', 'sunaengine_napi.node');
    if (!fs.existsSync(p)) throw new Error('[aether] N-API addon not built: python js/build_addon.py');
    addon = require(p);
  }
  return addon;
}

// ---- constants (src/engine.js:242-512) -----------------------------------
const ONE = 65536;
const H = 163840;
const CELL_SHIFT = 18;
const MAX_CELL_TOTAL = 1048576;
const MAX_POS = 134217728;
const VMAX = 65536;
const ITERS = 4;
const MAXNBR = 48;
const PARTICLE_WORDS = 8;
const DERIVED_WORDS = 16;
const FLAG_BURNING = 1, FLAG_DEAD = 2, FLAG_ANCHOR = 4, FLAG_SEAL = 128;
const FLAG_MASK = 7;
const SPAWN_FLAG_MASK = FLAG_BURNING | FLAG_ANCHOR;
const WELD_TAG_SHIFT = 8;
const TEMP_MIN = -33554432, TEMP_MAX = 536870912;
const AMBIENT = 20 * ONE;
const TABLE_HASH = '31ebf1f3e1b168a800f0f180007dd390d0dc182fd06e398adfeec08e8a62908b';
const MAT_WATER = 0;

const assert = (cond, msg) => { if (!cond) throw new Error('[aether] ' + msg); };

// ---- hashing (src/hash/statehash.js:33-54, restated) ----------------------
const rotl32 = (x, r) => (((x << (r & 31)) | (x >>> ((32 - r) & 31))) >>> 0);
function mmStep(h, k) {
  k = Math.imul(k, 0xcc9e2d51) >>> 0;
  k = rotl32(k, 15);
  k = Math.imul(k, 0x1b873593) >>> 0;
  h = (h ^ k) >>> 0;
  h = rotl32(h, 13);
  return (Math.imul(h, 5) + 0xe6546b64) >>> 0;
}
function fmix32(h) {
  h = (h ^ (h >>> 16)) >>> 0;
  h = Math.imul(h, 0x85ebca6b) >>> 0;
  h = (h ^ (h >>> 13)) >>> 0;
  h = Math.imul(h, 0xc2b2ae35) >>> 0;
  return (h ^ (h >>> 16)) >>> 0;
}
const SEED = [0x9e3779b1, 0x85ebca77, 0xc2b2ae3d, 0x27d4eb2f];
const SALT = [0x00000000, 0x51633e2d, 0x9c8f2a17, 0xd0e5b943];
function digestWords(u32, stride, hashed) {
  const acc = new Uint32Array(8);
  const n = hashed >>> 0, m = (hashed + 1) >>> 0;
  const count = (u32.length / stride) | 0;
  for (let i = 0; i < count; i++) {
    const h = SEED.slice();
    for (let w = 0; w < hashed; w++) {
      const k = u32[i * stride + w];
      for (let l = 0; l < 4; l++) h[l] = mmStep(h[l], (k ^ SALT[l]) >>> 0);
    }
    for (let l = 0; l < 4; l++) {
      acc[l] = (acc[l] + fmix32((h[l] ^ n) >>> 0)) >>> 0;
      acc[4 + l] = (acc[4 + l] + fmix32((mmStep(h[l], i >>> 0) ^ m) >>> 0)) >>> 0;
    }
  }
  return acc;
}
const hex8 = (a) => Array.from(a, (v) => (v >>> 0).toString(16).padStart(8, '0')).join('');

// ---- frozen data ----------------------------------------------------------
function loadLuts() {
  const buf = fs.readFileSync(path.join(ROOT, 'sunaengine_amd', 'data', 'luts.bin'));
  const img = new Int32Array(buf.buffer.slice(buf.byteOffset, buf.byteOffset + buf.byteLength));
  const h = hex8(digestWords(new Uint32Array(img.buffer), 1, 1));
  assert(h === TABLE_HASH, `TABLE_HASH mismatch: ${h}`);
  return img;
}
let MATS_JSON = null;
function mats() {
  if (!MATS_JSON) MATS_JSON = JSON.parse(fs.readFileSync(path.join(__dirname, 'build', 'mats.json'), 'utf8'));
  return MATS_JSON;
}
const F = (name) => mats().fields.indexOf(name);
function sealRow(row) {
  return row[F('bondK')] !== 0 && row[F('bondReform')] === 0 && row[F('solidMode')] !== 0 && row[F('phantom')] === 0;
}
function spawnWelds(row) { return row[F('bondGroup')] >= 0 && row[F('bondReform')] === 0; }

// ---- scenes (src/engine.js:2153-2169, 2396-2649, restated) ---------------
function xorshift32(seed) {
  let s = (seed >>> 0) || 0x2545f491;
  return () => {
    s ^= (s << 13); s >>>= 0;
    s ^= (s >>> 17);
    s ^= (s << 5); s >>>= 0;
    return s >>> 0;
  };
}
function emptyScene(n, nFluid, domWwu, domHwu) {
  return { n, nFluid, pos: new Int32Array(n * 2), vel: new Int32Array(n * 2), matId: new Uint32Array(n),
           domW: (domWwu * ONE) | 0, domH: (domHwu * ONE) | 0, cellShift: CELL_SHIFT };
}
function boundaryRing(domWwu, domHwu, wallRows) {
  const out = [];
  const half = ONE >> 1;
  for (let row = 0; row < 2; row++) {
    const y = domHwu * ONE - half - row * ONE;
    for (let c = 0; c < domWwu; c++) out.push([half + c * ONE, y]);
  }
  for (const xBase of [half, domWwu * ONE - half]) {
    for (let col = 0; col < 2; col++) {
      const x = xBase === half ? half + col * ONE : xBase - col * ONE;
      for (let r = 0; r < wallRows; r++) out.push([x, domHwu * ONE - half - (2 + r) * ONE]);
    }
  }
  return out;
}
function assemble(fluid, bnd, w, h) {
  const s = emptyScene(fluid.length + bnd.length, fluid.length, w, h);
  fluid.forEach((f, i) => { s.pos[2 * i] = f[0]; s.pos[2 * i + 1] = f[1]; s.matId[i] = f[2]; });
  bnd.forEach((b, k) => {
    const i = fluid.length + k;
    s.pos[2 * i] = b[0]; s.pos[2 * i + 1] = b[1]; s.matId[i] = MAT_WATER;
  });
  return s;
}
function makeScene(name, opts) {
  opts = opts || {};
  const DOM_W = 128, DOM_H = 72;
  if (name === 'dambreak_v1') {
    const rnd = xorshift32(0x1234567);
    const fluid = [];
    for (let gy = 0; gy < 60; gy++) {
      for (let gx = 0; gx < 60; gx++) {
        const jx = (rnd() % 4097) - 2048, jy = (rnd() % 4097) - 2048;
        fluid.push([3 * ONE + gx * ONE + jx, (DOM_H - 3) * ONE - gy * ONE + jy, MAT_WATER]);
      }
    }
    return assemble(fluid, boundaryRing(DOM_W, DOM_H, 60), DOM_W, DOM_H);
  }
  if (name === 'overdense_v1') {
    const SIDE = 32, SP = ONE >> 1, cx = 64 * ONE, cy = 40 * ONE;
    const fluid = [];
    for (let gy = 0; gy < SIDE; gy++)
      for (let gx = 0; gx < SIDE; gx++) fluid.push([cx + (gx - SIDE / 2) * SP, cy + (gy - SIDE / 2) * SP, MAT_WATER]);
    return assemble(fluid, boundaryRing(DOM_W, DOM_H, 60), DOM_W, DOM_H);
  }
  if (name === 'pool_v1') {
    const CW = opts.w !== undefined ? opts.w : 100, CH = opts.h !== undefined ? opts.h : 24;
    const mat = opts.mat !== undefined ? opts.mat : MAT_WATER;
    assert(13 + CW <= DOM_W - 2 && CH <= DOM_H - 5, `pool_v1 w=${CW} h=${CH} does not fit the domain`);
    const rnd = xorshift32(0x0be7a1);
    const fluid = [];
    for (let gy = 0; gy < CH; gy++) {
      for (let gx = 0; gx < CW; gx++) {
        const jx = (rnd() % 4097) - 2048, jy = (rnd() % 4097) - 2048;
        fluid.push([13 * ONE + gx * ONE + jx, (DOM_H - 3) * ONE - gy * ONE + jy, mat]);
      }
    }
    return assemble(fluid, boundaryRing(DOM_W, DOM_H, 60), DOM_W, DOM_H);
  }
  if (name === 'empty_v1') {
    const W = Math.max(8, opts.domWwu !== undefined ? opts.domWwu : DOM_W) | 0;
    const Hh = Math.max(16, opts.domHwu !== undefined ? opts.domHwu : DOM_H) | 0;
    return assemble([], boundaryRing(W, Hh, Hh - 12), W, Hh);
  }
  throw new Error(`[aether] unknown scene "${name}"`);
}

// ---- loadScene's validation and state image (src/engine.js:3704-3937) ----
function buildState(scene, maxParticles, maxCells) {
  const M = mats();
  const rows = M.rows;
  const matCount = rows.length;
  const { n, nFluid, pos, vel, matId } = scene;
  assert(Number.isInteger(n) && n >= 0, 'scene.n must be a non-negative integer');
  assert(n <= maxParticles, `scene.n = ${n} exceeds maxParticles = ${maxParticles}`);
  assert(nFluid >= 0 && nFluid <= n, 'scene.nFluid out of range');
  assert(pos.length === n * 2 && vel.length === n * 2 && matId.length === n, 'scene array lengths must be n*2 / n*2 / n');
  for (let i = 0; i < n; i++) {
    assert(matId[i] < matCount, `matId[${i}] = ${matId[i]} >= matCount ${matCount}`);
    const lo = i < nFluid ? 0 : -MAX_POS;
    assert(pos[2 * i] >= lo && pos[2 * i] <= MAX_POS && pos[2 * i + 1] >= lo && pos[2 * i + 1] <= MAX_POS,
      `pos[${i}] outside [${lo}, 2^27]`);
  }
  if (scene.flags) for (let i = 0; i < n; i++) {
    assert(Number.isInteger(scene.flags[i]) && (scene.flags[i] & ~FLAG_MASK) === 0, `flags[${i}] sets a reserved bit`);
  }
  if (scene.temp) for (let i = 0; i < n; i++) {
    assert(scene.temp[i] >= TEMP_MIN && scene.temp[i] <= TEMP_MAX, `temp[${i}] outside range`);
  }
  if (scene.fuel) for (let i = 0; i < n; i++) assert(scene.fuel[i] >= 0, `fuel[${i}] < 0`);
  const cellShift = scene.cellShift !== undefined ? scene.cellShift : CELL_SHIFT;
  assert(Number.isInteger(cellShift) && cellShift >= 1 && cellShift <= 27, `scene.cellShift = ${cellShift} outside the 1..27 contract`);
  assert(Math.pow(2, cellShift) >= H, `scene.cellShift = ${cellShift} gives cells smaller than the kernel radius`);
  const cellSize = Math.pow(2, cellShift);
  const gridW = Math.max(1, Math.ceil(scene.domW / cellSize)), gridH = Math.max(1, Math.ceil(scene.domH / cellSize));
  assert(gridW * gridH <= MAX_CELL_TOTAL, `cellTotal ${gridW * gridH} exceeds the 2-level scan limit`);
  assert(gridW * gridH <= maxCells, `cellTotal ${gridW * gridH} exceeds allocated maxCells ${maxCells}`);
  const st = new Int32Array(n * PARTICLE_WORDS);
  const stu = new Uint32Array(st.buffer);
  for (let i = 0; i < n; i++) {
    const m = matId[i] >>> 0, row = rows[m];
    const want = row[F('phantom')] !== 0 ? FLAG_DEAD : 0;
    const given = (scene.flags ? scene.flags[i] : 0) >>> 0;
    assert(!(given & FLAG_DEAD) || want !== 0, `scene.flags[${i}] sets FLAG_DEAD but matId ${m} is not phantom`);
    const o = i * PARTICLE_WORDS;
    st[o] = pos[2 * i]; st[o + 1] = pos[2 * i + 1]; st[o + 2] = vel[2 * i]; st[o + 3] = vel[2 * i + 1];
    stu[o + 4] = m;
    const tag = (i < nFluid && spawnWelds(row)) ? (1 << WELD_TAG_SHIFT) : 0;
    stu[o + 5] = (given | want | (sealRow(row) ? FLAG_SEAL : 0) | tag) >>> 0;
    st[o + 6] = scene.temp ? scene.temp[i] : AMBIENT;
    st[o + 7] = scene.fuel ? scene.fuel[i] : row[F('fuel0')];
  }
  return { n, nFluid, state: st, domW: scene.domW, domH: scene.domH, cellShift, gridW, gridH };
}

// ---- the engine -------------------------------------------------------------
class Engine {
  constructor(opts) {
    this.maxParticles = opts.maxParticles;
    this.maxCells = opts.maxCells;
    this.iters = opts.iters;
    const img = Int32Array.from(mats().image);
    this._h = native().create(opts.maxParticles, opts.maxCells, opts.maxScenes || 1, opts.iters,
      opts.streamCap || 4096, opts.device || 0, loadLuts(), img, opts.debug === false ? 0 : 1,
      opts.materials ? 1 : 0);
    this.n = 0; this.nFluid = 0; this._scene = null; this.substepCount = 0;
    // the per-slot RGBA8 colour sidecar (src/engine.js:4382-4402): not
    // simulation state, never hashed, kept on the host
    this._tints = new Uint32Array(Math.max(1, opts.maxParticles));
  }
  get matCount() { return mats().rows.length; }
  // stAcc (optional): n*2 words of carried surface-tension acceleration
  // (Derived words 8..9), for an exact restore.
  loadScene(scene, stAcc) {
    const im = buildState(scene, this.maxParticles, this.maxCells);
    if (stAcc) {
      assert(stAcc.length === im.n * 2, 'stAcc must hold n*2 words');
      im.stAcc = Int32Array.from(stAcc);
    }
    native().loadScenes(this._h, [im]);
    this.n = im.n; this.nFluid = im.nFluid; this._scene = scene;
    this.domW = im.domW; this.domH = im.domH; this.cellShift = im.cellShift;
    this.substepCount = 0;
    this._tints.fill(0);
    return this;
  }
  // `_substep` (src/engine.js:5825): substeps encoded since the load; it
  // decides the formation schedule and survives spawnFluid's reload
  get formationClock() { return native().substepCount(this._h, 0); }
  async readTints(count) {
    if (count === undefined) count = this.n;
    return this._tints.slice(0, Math.max(1, count));
  }
  setTints(first, words) {
    assert(Number.isInteger(first) && first >= 0 && first + words.length <= this.maxParticles,
      `setTints: [${first}, ${first + words.length}) is outside [0, ${this.maxParticles})`);
    if (!words.length) return 0;
    for (let i = 0; i < words.length; i++) this._tints[first + i] = words[i] >>> 0;
    return words.length;
  }
  // A device-resident push schedule (no reference counterpart: the
  // reference applies each push from the host): events {substep, cx, cy, r,
  // ix, iy, limit?} in arming order, each applied as applyImpulse
  // immediately before its substep, inside the captured substeps.
  setEvents(events) {
    const w = new Int32Array(Math.max(1, events.length) * 8);
    events.forEach((ev, k) => {
      w.set([0, ev.substep, ev.cx, ev.cy, ev.r, ev.ix, ev.iy, ev.limit === undefined ? 3000 : ev.limit], k * 8);
    });
    native().setEvents(this._h, events.length ? w : w.subarray(0, 0));
  }
  // Particles moved by scheduled events and applyImpulse since the load.
  pushHits() { return native().pushHits(this._h, 0); }
  step(count) {
    if (count === undefined) count = 1;
    assert(this.n > 0, 'step() called before loadScene()');
    assert(Number.isInteger(count) && count >= 0, 'step(count) needs a non-negative integer');
    if (count === 0) return this.substepCount;
    native().step(this._h, count);
    this.substepCount += count;
    return this.substepCount;
  }
  async digest() {
    const acc = native().readHash(this._h, 0);
    return { acc, hex: hex8(acc), setHex: hex8(acc.subarray(0, 4)), slotHex: hex8(acc.subarray(4, 8)) };
  }
  async chainHex() { return hex8(native().readChain(this._h, 0)); }
  hashStream() { return native().readHashStream(this._h, 0, native().substepCount(this._h, 0) + 1); }
  resetStream() { native().resetStream(this._h); }
  async readState() { return native().readState(this._h, 0, this.n); }
  async readDerived() { return native().readDerived(this._h, 0, this.n); }
  async readBond() { return native().readBond(this._h, 0, this.n); }
  // The buffer writes of restore(snap) (src/demo_boot.js:308-321): state and
  // derived words of the loaded scene, slot order.
  writeStateDerived(state, derived) {
    native().writeStateDerived(this._h, 0, Int32Array.from(state), derived ? Int32Array.from(derived) : null);
  }
  // Tune the launch shape on this world (results unchanged; see sn_autotune).
  autotune(trialSubsteps, lookahead) { return native().autotune(this._h, trialSubsteps || 16, lookahead || 0); }
  async debugCounters() { return native().debugCounters(this._h, 0); }
  async checkErrors() { native().sync(this._h); }
  // The exact current state is what applyImpulse selects on, so the mirror
  // is always in sync; kept for API shape (src/engine.js:6060).
  async syncMirror() { native().sync(this._h); return this.readState(); }
  applyImpulse(cx, cy, radius, ix, iy, limit) {
    if (limit === undefined) limit = 3000;
    return native().applyImpulse(this._h, 0, cx, cy, radius, ix, iy, limit);
  }
  spawnFluid(list, opts) {
    opts = opts || {};
    const k = list.length;
    if (k === 0) return this.n;
    const rows = mats().rows;
    const fl = new Int32Array(k * 8), flu = new Uint32Array(fl.buffer);
    list.forEach((p, i) => {
      const m = Math.min(rows.length - 1, Math.max(0, p.mat | 0)) >>> 0;
      const row = rows[m], o = i * 8;
      fl[o] = p.x | 0; fl[o + 1] = p.y | 0; fl[o + 2] = p.vx | 0; fl[o + 3] = p.vy | 0;
      flu[o + 4] = m;
      flu[o + 5] = (((p.flags !== undefined ? p.flags : 0) & SPAWN_FLAG_MASK) |
        (row[F('phantom')] !== 0 ? FLAG_DEAD : 0) | (sealRow(row) ? FLAG_SEAL : 0)) >>> 0;
      fl[o + 6] = Math.max(TEMP_MIN, Math.min(TEMP_MAX, p.temp !== undefined ? p.temp : row[F('spawnTemp')])) | 0;
      fl[o + 7] = Math.max(0, p.fuel !== undefined ? p.fuel : row[F('fuel0')]) | 0;
    });
    // the engine shifts its own boundary block and assigns the spawn's weld
    // group (src/engine.js:5933-5971)
    const n = native().spawnFluid(this._h, fl, (opts.noWeld ? 1 : 0) | (opts.ownGroup === true ? 2 : 0));
    if (n !== this.n) {
      // the colour block moves with the matter: the new slots take their
      // tints, the shifted boundary block is cleared (src/engine.js:5993-6004)
      const nB = this.n - this.nFluid;
      for (let i = 0; i < k; i++) this._tints[this.nFluid + i] = (list[i].tint !== undefined ? list[i].tint : 0) >>> 0;
      this._tints.fill(0, this.nFluid + k, this.nFluid + k + nB);
      this.nFluid += k; this.n = n;
    }
    return this.n;
  }
  async assertSpecValidRegime(where) {
    const dbg = await this.debugCounters();
    const names = ['mul/shift saturation', 'divide saturation or bad shift', 'divide by zero', 'sqrt of negative',
      'rsqrt of non-positive', 'add/sub/neg saturation', 'neighbour-list truncation', 'cell/scan capacity overflow',
      'bond row saturation'];
    const hit = Array.from(dbg).map((v, i) => (v ? `${names[i] || `RESERVED dbg[${i}]`} (dbg[${i}]=${v})` : null))
      .filter(Boolean);
    assert(hit.length === 0, `${where || 'run'}: simulation left the spec-valid regime — ${hit.join('; ')}`);
    return true;
  }
  destroy() { if (this._h) { native().destroy(this._h); this._h = null; } }
}

async function createEngine(opts) {
  opts = opts || {};
  const iters = opts.iters !== undefined ? opts.iters : ITERS;
  assert(Number.isInteger(iters) && iters >= 1 && iters <= 16, `iters = ${iters} must be an integer in 1..16`);
  return new Engine({ maxParticles: opts.maxParticles || 20000, maxCells: opts.maxCells || 65536, iters,
    maxScenes: 1, streamCap: opts.streamCap || 4096, device: opts.device || 0, debug: opts.debug !== false,
    materials: !!opts.materials });
}

module.exports = {
  createEngine, Engine, makeScene, buildState, digestWords, hex8, mmStep, fmix32, xorshift32, boundaryRing,
  native, loadLuts, matsImage: () => Int32Array.from(mats().image),
  ONE, H, VMAX, MAXNBR, ITERS, PARTICLE_WORDS, DERIVED_WORDS, HASHED_WORDS: 8, MAT_WATER, TABLE_HASH,
  FLAG_MASK, AMBIENT,
};

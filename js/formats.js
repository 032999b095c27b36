'use strict';
// Wire formats of the deterministic harness.
//
//  * .suna: one world as JSON, with the particle state and the tint sidecar
//    as base64 little-endian word arrays (src/suna.js:46-124).
//  * .sunar: a replay journal, i.e. an initial .suna world plus
//    frame/order-indexed input events and timing (src/sunar.js:27-92).
//
// The JSON layout, the validation rules and the thrown texts are the
// reference's interface and are kept; the code is this host's.  Node 12
// parses ES2019, so no `??` or `?.`.

const FORMAT = {
  suna: { magic: 'SunaBox Scene', version: 1, maxParticles: 131072 },
  sunar: { magic: 'SunaBox Replay', version: 1 },
};
const I32_MIN = -2147483648, I32_MAX = 2147483647;

// The range check both codecs apply to every integer field.
function checkInt(x, label, lo, hi) {
  if (!Number.isInteger(x) || x < lo || x > hi) throw new Error(`${label} is outside ${lo}..${hi}`);
  return x;
}

// Typed word arrays <-> base64 of their little-endian bytes.
const words = {
  encode(a) { return Buffer.from(a.buffer, a.byteOffset, a.byteLength).toString('base64'); },
  decode(text, Ctor, count, label) {
    const bytes = Buffer.from(text || '', 'base64');
    if (bytes.byteLength !== count * 4) throw new Error(`${label} has ${bytes.byteLength} bytes; expected ${count * 4}`);
    const own = new ArrayBuffer(bytes.byteLength);      // aligned, owned copy
    new Uint8Array(own).set(bytes);
    return new Ctor(own);
  },
};

// Parse a document and check its magic and version.
function openDocument(source, fmt, ext) {
  const text = typeof source === 'string' ? source : Buffer.from(source).toString('utf8');
  let doc;
  try {
    doc = JSON.parse(text);
  } catch (err) {
    throw new Error(`not valid JSON: ${err.message}`);
  }
  if (!doc || doc.magic !== fmt.magic) throw new Error(`not a ${fmt.magic} file`);
  if (doc.formatVersion !== fmt.version) {
    throw new Error(`unsupported ${ext} formatVersion ${doc.formatVersion}; this build reads 1`);
  }
  return doc;
}

// Journal events must be in stable (frame, order) sequence, each typed.
function checkJournal(events) {
  let last = [-1, -1];
  for (let i = 0; i < events.length; i++) {
    const e = events[i] || {};
    const at = [checkInt(e.frame, `events[${i}].frame`, 0, I32_MAX), checkInt(e.order, `events[${i}].order`, 0, I32_MAX)];
    if (at[0] < last[0] || (at[0] === last[0] && at[1] <= last[1])) {
      throw new Error(`events[${i}] is not in stable frame/order sequence`);
    }
    if (typeof e.type !== 'string' || e.type.length === 0) throw new Error(`events[${i}].type is empty`);
    last = at;
  }
}

// ---- .suna -----------------------------------------------------------------
function encodeSuna(payload) {
  const state = payload.state, tints = payload.tints;
  if (!(state instanceof Int32Array)) throw new Error('Suna save needs an Int32Array state');
  if (!(tints instanceof Uint32Array)) throw new Error('Suna save needs a Uint32Array tint buffer');
  const scene = payload.scene || {};
  const n = checkInt(scene.n, 'scene.n', 0, FORMAT.suna.maxParticles);
  checkInt(scene.nFluid, 'scene.nFluid', 0, n);
  if (state.length !== n * 8) throw new Error(`state has ${state.length} words; expected ${n * 8}`);
  if (tints.length !== n) throw new Error(`tints has ${tints.length} words; expected ${n}`);
  const body = Object.assign({}, scene, {
    state: { encoding: 'i32le/base64', words: state.length, data: words.encode(state) },
    tints: { encoding: 'u32le/base64', words: tints.length, data: words.encode(tints) },
  });
  const doc = {
    magic: FORMAT.suna.magic, formatVersion: FORMAT.suna.version, app: payload.app, scene: body,
    materials: payload.materials || { events: [] }, bodies: payload.bodies || [], drivers: payload.drivers || [],
    devices: payload.devices || [], settings: payload.settings || {},
  };
  return JSON.stringify(doc, null, 2) + '\n';
}

function decodeSuna(source) {
  const doc = openDocument(source, FORMAT.suna, '.suna');
  const s = doc.scene || {};
  const n = checkInt(s.n, 'scene.n', 0, FORMAT.suna.maxParticles);
  checkInt(s.nFluid, 'scene.nFluid', 0, n);
  checkInt(s.domW, 'scene.domW', 1, I32_MAX);
  checkInt(s.domH, 'scene.domH', 1, I32_MAX);
  checkInt(s.cellShift, 'scene.cellShift', 1, 27);
  const state = words.decode(s.state && s.state.data, Int32Array, n * 8, 'scene.state');
  const tints = words.decode(s.tints && s.tints.data, Uint32Array, n, 'scene.tints');
  const scene = {};
  for (const k of Object.keys(s)) if (k !== 'state' && k !== 'tints') scene[k] = s[k];
  return Object.assign({}, doc, { scene, state, tints });
}

// ---- .sunar ----------------------------------------------------------------
function encodeSunar(replay) {
  if (!replay || !replay.initial) throw new Error('Suna replay needs an initial scene');
  const events = Array.isArray(replay.events) ? replay.events : [];
  checkJournal(events);
  const initial = JSON.parse(encodeSuna(replay.initial));
  const timing = {
    substepsPerFrame: checkInt(replay.substepsPerFrame === undefined ? 2 : replay.substepsPerFrame,
      'substepsPerFrame', 1, 32),
    durationFrames: checkInt(replay.durationFrames || 0, 'durationFrames', 0, I32_MAX),
    durationSubsteps: checkInt(replay.durationSubsteps || 0, 'durationSubsteps', 0, I32_MAX),
  };
  const doc = { magic: FORMAT.sunar.magic, formatVersion: FORMAT.sunar.version, app: replay.app || initial.app,
    timing, initial, events, final: replay.final || null };
  return JSON.stringify(doc, null, 2) + '\n';
}

function decodeSunar(source) {
  const doc = openDocument(source, FORMAT.sunar, '.sunar');
  const t = doc.timing || {};
  checkInt(t.substepsPerFrame, 'timing.substepsPerFrame', 1, 32);
  checkInt(t.durationFrames, 'timing.durationFrames', 0, I32_MAX);
  checkInt(t.durationSubsteps, 'timing.durationSubsteps', 0, I32_MAX);
  const events = Array.isArray(doc.events) ? doc.events : [];
  checkJournal(events);
  return {
    app: doc.app, substepsPerFrame: t.substepsPerFrame, durationFrames: t.durationFrames,
    durationSubsteps: t.durationSubsteps, initial: decodeSuna(JSON.stringify(doc.initial)), events,
    final: doc.final || null,
  };
}

module.exports = { encodeSuna, decodeSuna, encodeSunar, decodeSunar, FORMAT };

'use strict';
// CPU-side checks of the Node host (no device calls): frozen tables, scene
// builders and the state digest against values the Python side computes,
// the .suna / .sunar codecs, and loadScene's contract errors.
// Usage: node host_cpu.js '<json of expected values>'
const assert = require('assert');
const eng = require('../../js/sunaengine');
const fm = require('../../js/formats');
const hz = require('../../js/harness');

const want = JSON.parse(process.argv[2]);
let checks = 0;
const ok = (c, msg) => { assert.ok(c, msg); checks++; };

// frozen kernel tables: loadLuts checks TABLE_HASH itself
const luts = (() => {
  const fs = require('fs'), path = require('path');
  return fs.readFileSync(path.join(__dirname, '..', '..', 'sunaengine_amd', 'data', 'luts.bin'));
})();
ok(luts.length === 16010 * 4, 'LUT image size');
ok(eng.hex8(eng.digestWords(new Uint32Array(luts.buffer, luts.byteOffset, 16010), 1, 1)) === eng.TABLE_HASH,
  'TABLE_HASH of the frozen image');

// scene builders and the state digest equal the Python restatement's
for (const name of Object.keys(want.scenes)) {
  const w = want.scenes[name];
  const sc = name === 'pool_v1' ? eng.makeScene(name, { w: 60, h: 20 }) : eng.makeScene(name);
  ok(sc.n === w.n && sc.nFluid === w.nFluid, `${name}: n/nFluid`);
  const im = eng.buildState(sc, 1 << 20, 1 << 20);
  const h = eng.hex8(eng.digestWords(new Uint32Array(im.state.buffer), 8, 8));
  ok(h === w.digest, `${name}: digestState of the loaded image ${h} != ${w.digest}`);
}

// loadScene's contract errors carry the reference's text
const dam = eng.makeScene('dambreak_v1');
assert.throws(() => eng.buildState(dam, 100, 1 << 20), /\[aether\] scene\.n = 4096 exceeds maxParticles = 100/);
assert.throws(() => eng.buildState(dam, 1 << 20, 10), /exceeds allocated maxCells/);
checks += 2;

// .suna round trip
const st = new Int32Array(8 * 3).map((_, i) => (i * 7919) | 0);
const payload = { app: { name: 'test' }, scene: { n: 3, nFluid: 2, domW: 8388608, domH: 4718592, cellShift: 18 },
  state: st, tints: new Uint32Array([1, 2, 3]) };
const text = fm.encodeSuna(payload);
const back = fm.decodeSuna(text);
ok(back.state.length === 24 && back.state.every((v, i) => v === st[i]), '.suna state round trip');
ok(back.tints[2] === 3 && back.scene.n === 3, '.suna scene fields');
assert.throws(() => fm.decodeSuna('{"magic":"nope"}'), /not a SunaBox Scene file/);
assert.throws(() => fm.decodeSuna(text.replace('"formatVersion": 1', '"formatVersion": 2')), /unsupported \.suna formatVersion/);
checks += 2;

// .sunar round trip and ordering rules
const replay = { initial: payload, substepsPerFrame: 1, durationFrames: 9, durationSubsteps: 9,
  events: [{ frame: 2, order: 0, type: 'inject', particles: [{ x: 1, y: 2, vx: 0, vy: 0 }] },
    { frame: 2, order: 1, type: 'push', x: 5, y: 6, r: 7, ix: 1, iy: 0 }] };
const rt = fm.decodeSunar(fm.encodeSunar(replay));
ok(rt.events.length === 2 && rt.events[1].type === 'push' && rt.initial.state[5] === st[5], '.sunar round trip');
assert.throws(() => fm.encodeSunar(Object.assign({}, replay, { events: [replay.events[1], replay.events[0]] })),
  /not in stable frame\/order sequence/);
checks++;

// .suna field checks (the reference's texts)
assert.throws(() => fm.encodeSuna(Object.assign({}, payload, { tints: new Uint32Array(2) })), /tints has 2 words; expected 3/);
assert.throws(() => fm.decodeSuna(text.replace('"cellShift": 18', '"cellShift": 28')), /scene\.cellShift is outside 1\.\.27/);
checks += 2;

// the harness's input timeline: (substep, arming order), ranges taken in order
const tl = new hz.Timeline();
tl.arm(5, 'push', 'a'); tl.arm(2, 'inject', 'b'); tl.arm(5, 'inject', 'c'); tl.arm(5, 'push', 'd'); tl.arm(9, 'push', 'e');
ok(tl.nextAfter(2, 'inject') === 5 && tl.nextAfter(5, 'inject') === Infinity, 'timeline nextAfter');
ok(tl.takeRange(3, 9, 'push').map((it) => it.data).join('') === 'ad', 'timeline takeRange keeps arming order');
ok(tl.items.map((it) => it.data).join('') === 'bce', 'timeline keeps the rest sorted');

// droplet helper: w x h grid centred on (cx, cy), one wu apart
const d = hz.makeDroplet(hz.wu(10), hz.wu(20), 3, 2, 5, -5);
ok(d.length === 6 && d[0].x === hz.wu(9) && d[5].y === hz.wu(20) + (eng.ONE >> 1), 'makeDroplet layout');

console.log(`OK ${checks} checks`);

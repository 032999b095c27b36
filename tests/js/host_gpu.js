'use strict';
// The Node host on the GPU through the N-API addon: the reference's frozen
// dambreak_v1 gate, the water harness's record -> .sunar -> replay round trip
// (same final digest), and exact snapshot/restore.
// Usage: node host_gpu.js '<json: golden checkpoints>'
const assert = require('assert');
const eng = require('../../js/sunaengine');
const rp = require('../../js/replay');
const golden = JSON.parse(process.argv[2]);

async function main() {
  const e = await eng.createEngine({ maxParticles: 8192, maxCells: 65536, iters: 4 });
  e.loadScene(eng.makeScene('dambreak_v1'));
  let done = 0;
  for (const c of [1, 10, 100]) {
    e.step(c - done);
    done = c;
    const d = await e.digest();
    assert.strictEqual(d.hex, golden[String(c)], `dambreak_v1 checkpoint ${c}`);
    console.log(`dambreak_v1 checkpoint ${c} ok`);
  }
  await e.assertSpecValidRegime('dambreak_v1');
  e.destroy();

  // record a session with injections and pushes, replay it from .sunar text
  const demo = await rp.bootWaterDemo({ scene: 'pool', poolW: 40, poolH: 12, maxParticles: 3000 });
  demo.step(7);
  await demo.startRecording();
  const s0 = demo.substep;
  demo.injectWater({ substep: s0 + 5, particles: rp.makeDroplet(rp.wu(60), rp.wu(30), 3, 3, 0, 2000) });
  demo.pushAt({ substep: s0 + 12, x: rp.wu(40), y: rp.wu(64), r: rp.wu(6), ix: 30000, iy: -20000 });
  demo.injectWater({ substep: s0 + 20, particles: rp.makeDroplet(rp.wu(30), rp.wu(28), 2, 2, 1000, 0) });
  await demo.stepThrough(40);
  const { text } = await demo.stopRecording();
  const finalDigest = await demo.digest();
  const nAfter = demo.n;
  demo.destroy();
  const rep = await rp.replayFromSunar(text);
  await rep.stepThrough(40);
  assert.strictEqual(rep.n, nAfter, 'replay particle count');
  assert.strictEqual(await rep.digest(), finalDigest, 'replay reproduces the recorded run');
  console.log('record/replay ok');

  // exact snapshot / restore
  rep.step(5);
  const snap = await rep.snapshot({ exact: true });
  rep.step(30);
  const a = await rep.digest();
  rep.restore(snap);
  rep.step(30);
  assert.strictEqual(await rep.digest(), a, 'restore(snapshot) continues identically');
  rep.destroy();

  // the engine-level restore writes (writeStateDerived) rejoin a timeline too
  const e1 = await eng.createEngine({ maxParticles: 8192, maxCells: 65536, iters: 4 });
  e1.loadScene(eng.makeScene('dambreak_v1'));
  e1.step(9);
  const st = await e1.readState(), dv = await e1.readDerived();
  e1.step(20);
  const e2 = await eng.createEngine({ maxParticles: 8192, maxCells: 65536, iters: 4 });
  e2.loadScene(eng.makeScene('dambreak_v1'));
  e2.writeStateDerived(st, dv);
  e2.step(20);
  assert.strictEqual((await e2.digest()).hex, (await e1.digest()).hex, 'writeStateDerived continues identically');
  e1.destroy();
  e2.destroy();
  console.log('OK node host');
}
main().catch((err) => { console.error(err && err.stack ? err.stack : err); process.exit(1); });

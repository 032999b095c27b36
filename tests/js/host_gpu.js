'use strict';
// The Node host on the GPU through the N-API addon: the reference's frozen
// dambreak_v1 gate, the water harness's record -> .sunar -> replay round trip
// (same final digest), and exact snapshot/restore.
// Usage: node host_gpu.js '<json: golden checkpoints>'
const assert = require('assert');
const eng = require('../../js/sunaengine');
const hz = require('../../js/harness');
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

  // record a session with injections and pushes (one push on the substep of
  // an injection, two on one substep, the rest run from the device schedule),
  // replay it from .sunar text
  const demo = await hz.bootWaterDemo({ scene: 'pool', poolW: 40, poolH: 12, maxParticles: 3000 });
  demo.step(7);
  await demo.startRecording();
  const s0 = demo.substep;
  demo.injectWater({ substep: s0 + 5, particles: hz.makeDroplet(hz.wu(60), hz.wu(30), 3, 3, 0, 2000) });
  demo.pushAt({ substep: s0 + 12, x: hz.wu(40), y: hz.wu(64), r: hz.wu(6), ix: 30000, iy: -20000 });
  demo.pushAt({ substep: s0 + 12, x: hz.wu(42), y: hz.wu(64), r: hz.wu(5), ix: -65536, iy: 5000 });
  demo.pushAt({ substep: s0 + 20, x: hz.wu(30), y: hz.wu(66), r: hz.wu(8), ix: 20000, iy: 20000 });
  demo.injectWater({ substep: s0 + 20, particles: hz.makeDroplet(hz.wu(30), hz.wu(28), 2, 2, 1000, 0) });
  demo.pushAt({ substep: s0 + 31, x: hz.wu(50), y: hz.wu(66), r: hz.wu(9), ix: -10000, iy: -40000 });
  assert.throws(() => demo.step(40), /push armed at substep .*: use stepThrough\(\)/);
  await demo.stepThrough(40);
  const hits = demo.pushHits;
  assert.ok(hits > 0, 'the pushes moved particles');
  const { text } = await demo.stopRecording();
  const finalDigest = await demo.digest();
  const nAfter = demo.n;
  demo.destroy();
  const rep = await hz.replayFromSunar(text);
  await rep.stepThrough(40);
  assert.strictEqual(rep.n, nAfter, 'replay particle count');
  assert.strictEqual(await rep.digest(), finalDigest, 'replay reproduces the recorded run');
  assert.strictEqual(rep.pushHits, hits, 'replay moves the same particles');
  console.log('record/replay ok');

  // the same inputs applied one substep at a time from the host (the
  // reference's own path: sync, applyImpulse in arming order, then the
  // injection) give the same world
  const e3 = await eng.createEngine({ maxParticles: 3000 });
  const doc = rep.replayDoc;
  const seedDemo = new hz.WaterDemo(e3, 3000);
  seedDemo._loadPayload(doc);
  let hostHits = 0;
  for (let s = 0; s < 40; s++) {
    const due = doc.events.filter((ev) => ev.frame === s);
    for (const ev of due) if (ev.type === 'push') hostHits += e3.applyImpulse(ev.x, ev.y, ev.r, ev.ix, ev.iy);
    const parts = [];
    for (const ev of due) if (ev.type === 'inject') for (const p of ev.particles) parts.push(Object.assign({}, p, { mat: eng.MAT_WATER }));
    if (parts.length) e3.spawnFluid(parts, { noWeld: true });
    e3.step(1);
  }
  assert.strictEqual((await e3.digest()).hex, finalDigest, 'host-applied inputs give the same world');
  assert.strictEqual(hostHits, hits, 'host-applied pushes move the same particles');
  e3.destroy();
  console.log('device schedule = host pushes ok');

  // exact snapshot / restore; the snapshot carries tints and both clocks
  rep.step(5);
  rep.engine.setTints(3, [0x11223344, 0x55667788]);
  const snap = await rep.snapshot({ exact: true });
  assert.strictEqual(snap.scene.substepCount, 45, 'snapshot substepCount');
  assert.strictEqual(snap.scene.formationClock, 45, 'snapshot formationClock');
  assert.strictEqual(snap.tints[4], 0x55667788, 'snapshot tints');
  rep.step(30);
  const a = await rep.digest();
  rep.restore(snap);
  assert.strictEqual((await rep.engine.readTints())[3], 0x11223344, 'restore puts the tints back');
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

'use strict';
// The Node host on a material scene: loadScene -> step -> per-substep hash
// stream and bond rows, printed as JSON for the Python side to compare with
// the oracle.  Usage: node materials_gpu.js <scene.json> <substeps>
const fs = require('fs');
const eng = require('../../js/sunaengine');

async function main() {
  const sc = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
  const steps = parseInt(process.argv[3], 10);
  const scene = {
    n: sc.n, nFluid: sc.nFluid, domW: sc.domW, domH: sc.domH,
    pos: Int32Array.from(sc.pos), vel: Int32Array.from(sc.vel), matId: Uint32Array.from(sc.matId),
    flags: sc.flags ? Uint32Array.from(sc.flags) : undefined,
    temp: sc.temp ? Int32Array.from(sc.temp) : undefined,
  };
  const e = await eng.createEngine({ maxParticles: sc.n, maxCells: 65536, iters: 4, materials: !!sc.materials });
  e.loadScene(scene);
  e.step(2);
  // tuning the launch shape on the world leaves every result unchanged
  const tune = e.autotune(4, 2);
  if (!tune.chosen || tune.trials.length !== 12) throw new Error('autotune report');
  e.step(steps - 2);
  const stream = e.hashStream();
  const hex = [];
  for (let s = 0; s < stream.length / 8; s++) {
    let h = '';
    for (let l = 0; l < 8; l++) h += ('00000000' + stream[s * 8 + l].toString(16)).slice(-8);
    hex.push(h);
  }
  const bond = await e.readBond();
  let bonds = 0, sum = 0;
  for (let k = 0; k < bond.length; k++) if (bond[k] !== 0) { bonds++; sum = (sum + bond[k] * (k + 1)) >>> 0; }
  console.log(JSON.stringify({ stream: hex, bonds, bondSum: sum }));
  e.destroy();
}
main().catch((err) => { console.error(err && err.stack ? err.stack : err); process.exit(1); });

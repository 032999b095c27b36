'use strict';
// The JavaScript serving loop on the GPU: every job's per-substep stream is
// the same in the continuous-batching mix (sharing worlds, admitted and
// retired between rounds, partitions reused, stripped build) as alone
// (debug build), for 2D and 3D jobs.  Usage: node serve_gpu.js
const serve = require('../../js/serve.js');

const ONE = 65536;
const specs = [['2d', 3000, 70], ['3d', 1500, 45], ['2d', 6000, 33], ['2d', 2500, 90],
  ['3d', 2700, 60], ['2d', 4200, 21], ['3d', 1000, 38]];
const jobs = specs.map(([kind, n, steps], k) => serve.makeJob(k, 1000 + 17 * k, kind, n, steps));
for (const j of jobs) {   // a strong stir at substep 3 of every job, beside its seeded ones
  if (j.kind === '2d') {
    j.events.push([3, Math.floor(j.scene.domW / 2), Math.floor(j.scene.domH * 3 / 4), 8 * ONE, ONE, -30000, 3000]);
  } else {
    const w = j.scene.gx * 2 * ONE;
    j.events.push([3, Math.floor(w / 2), Math.floor(w / 4), Math.floor(w / 2), 8 * ONE, 40000, 20000, -10000, 3000]);
  }
  j.events.sort((a, b) => a[0] - b[0]);
}
const srv = new serve.Server({ classes2d: [[7000, 2048, 2]], classes3d: [[3000, 4096, 2]], chunk: 16, streamCap: 256 });
const st = srv.run(jobs, { realtime: false });
srv.close();
console.log(`served ${jobs.length} jobs in ${st.rounds['2d']} + ${st.rounds['3d']} rounds`);
let bad = 0;
for (const j of jobs) {
  const ref = serve.runAlone(j);
  const same = j.stream && j.stream.length === ref.length && j.stream.every((v, i) => v === ref[i]);
  console.log(`job ${j.jid} ${j.kind} n=${j.n} x${j.substeps} events=${j.events.length}: ${same ? 'equal' : 'DIFFERENT'}`);
  if (!same) bad++;
}
if (bad) { console.log(`${bad} jobs differ`); process.exit(1); }
console.log('OK js serving mix = alone');

'use strict';
// Serving benchmark of the JavaScript host: a Poisson stream of replay/verify
// jobs per arrival rate, served by js/serve.js; prints one JSON object with
// the aggregate particle-substeps/s and each rate's p50/p99 latency.
// Usage: node serve_bench.js <rates, e.g. 10,100,500> <jobs per rate> <max substeps>
const serve = require('./serve.js');

const rates = (process.argv[2] || '10,100,500').split(',').map(Number);
const perRate = Number(process.argv[3] || 12);
const maxSteps = Number(process.argv[4] || 5000);
const srv = new serve.Server();
const per = [];
let work = 0, busy = 0;
rates.forEach((rate, ri) => {
  const jobs = serve.poissonJobs(perRate, rate, (0x5E7E + 101 * ri) >>> 0, { stepsHi: maxSteps });
  const st = srv.run(jobs);
  const s = serve.summarize(jobs, st.wallS);
  s.rateJobsPerS = rate;
  per.push(s);
  work += s.particleSubsteps;
  busy += st.wallS;
  process.stderr.write(`# js host, rate ${rate}/s: ${s.jobs} jobs, ${s.throughput.toExponential(3)} ps/s, ` +
                       `p50 ${s.p50LatencyS.toFixed(2)}s p99 ${s.p99LatencyS.toFixed(2)}s\n`);
});
srv.close();
console.log(JSON.stringify({ value: work / busy, rates: per }));

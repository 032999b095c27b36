'use strict';
// Kernel, wall and ball tables of the 3D solver, built in IEEE float64: the
// JavaScript form of sunaengine_amd/tables3d.py (which restates buildLUTs /
// buildBallLUT / matWords of 3d/app.js:85-231).  Only + - * / and sqrt are
// used, which IEEE-754 rounds exactly, so every host builds the same
// integers; tests/js/host_cpu.js checks this image word for word against the
// Python one.
//
// Image layout (int32 words, 3d/sim3d.wgsl:63-76): W[2049] | G[2049] |
// planeW[513] | planeG[513] | ballW[1024] | ballG[1024] | 6 materials x 4.

const LUT_N = 2049, PN = 513, BALL_N = 1024;
const RHO0 = 1048576, H_WU = 2.0, SPACING = 1.0;
const OFF_W = 0, OFF_G = LUT_N, OFF_P = 2 * LUT_N, OFF_PG = 2 * LUT_N + PN;
const OFF_BW = 2 * LUT_N + 2 * PN, OFF_BG = OFF_BW + BALL_N, OFF_MAT = OFF_BG + BALL_N;
const BALL_R_DEFAULT = 3.4;
// 3d/app.js:180-199: grav, rhoW, damp, f0 (f0 is used by the renderer only)
const MATERIALS3D = [
  [180, 0, 65340, 20], [110, 30, 65180, 35], [440, 0, 65420, 700],
  [-46, 120, 64700, 10], [88, 210, 65330, 60], [92, 50, 62900, 45],
];

function buildLuts3d(ballR) {
  if (ballR === undefined) ballR = BALL_R_DEFAULT;
  const h = H_WU, h2 = h * h;
  const rawW = new Float64Array(LUT_N), rawG = new Float64Array(LUT_N);
  for (let ui = 0; ui < LUT_N; ui++) {
    const r2 = ui / 512;
    if (r2 >= h2) continue;
    const t = h2 - r2;
    rawW[ui] = t * t * t;
    const r = Math.sqrt(r2), hr = h - r;
    rawG[ui] = r > 1e-6 ? (hr * hr) / r : 0;
  }
  for (let ui = 0; ui < LUT_N; ui++) if (rawG[ui] > 0) { rawG[0] = rawG[ui]; break; }
  let sumRaw = 0;
  const rad = Math.ceil(h / SPACING);
  for (let i = -rad; i <= rad; i++)
    for (let j = -rad; j <= rad; j++)
      for (let k = -rad; k <= rad; k++) {
        const r2 = (i * i + j * j + k * k) * SPACING * SPACING;
        if (r2 >= h2) continue;
        const t = h2 - r2;
        sumRaw += t * t * t;
      }
  const SW = RHO0 / sumRaw, SG = SW * 6 * (h2 - 1) * (h2 - 1);
  const planeW = new Array(PN).fill(0), planeG = new Array(PN).fill(0);
  const lat = Math.ceil(h / SPACING);
  for (let b = 0; b < PN; b++) {
    const dist = b / 256;
    let sw = 0, sg = 0;
    for (let k = 0; ; k++) {
      const pd = dist + k * SPACING;
      if (pd >= h) break;
      for (let i = -lat; i <= lat; i++)
        for (let j = -lat; j <= lat; j++) {
          const r2 = pd * pd + (i * SPACING) * (i * SPACING) + (j * SPACING) * (j * SPACING);
          if (r2 >= h2) continue;
          const t = h2 - r2;
          sw += t * t * t;
          sg += pd * t * t;
        }
    }
    planeW[b] = Math.round(sw * SW);
    planeG[b] = Math.round(6 * sg * SW);
  }
  const img = new Int32Array(OFF_MAT + MATERIALS3D.length * 4);
  for (let ui = 0; ui < LUT_N; ui++) {
    img[OFF_W + ui] = Math.round(rawW[ui] * SW);
    img[OFF_G + ui] = Math.round(rawG[ui] * SG);
  }
  for (let b = 0; b < PN; b++) { img[OFF_P + b] = planeW[b]; img[OFF_PG + b] = planeG[b]; }
  const at = (tab, d) => tab[Math.max(0, Math.min(512, Math.round(d * 256)))];
  for (let b = 0; b < BALL_N; b++) {
    const r = Math.sqrt(b * 4096) / 256, depth = r - ballR;
    if (depth >= h) continue;
    const d = Math.max(depth, 0);
    img[OFF_BW + b] = at(planeW, d);
    img[OFF_BG + b] = r > 1e-6 ? Math.min(Math.round(at(planeG, d) / r), 1 << 24) : 0;
  }
  MATERIALS3D.forEach((m, i) => { for (let w = 0; w < 4; w++) img[OFF_MAT + i * 4 + w] = m[w]; });
  return img;
}

module.exports = { buildLuts3d, BALL_R_DEFAULT, LUT_WORDS: OFF_MAT + MATERIALS3D.length * 4 };

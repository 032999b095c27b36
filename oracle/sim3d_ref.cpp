// ORACLE — test infrastructure only.  Scalar CPU restatement of the
// reference's 3D solver (3d/sim3d.wgsl) driven in the host's pass order
// (3d/app.js:673-697), plus the reference's striped FNV digest
// (3d/sim3d.wgsl:571-590 folded as 3d/app.js:717-743).
//
// The 3D solver has no MurmurHash state hash of its own; the per-substep
// H_set / H_slot here is digestState (src/hash/statehash.js:80-159) over each
// particle's record with stride 16 words and hashed words p.xyzw, v.xyzw —
// the same definition the GPU path uses.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "fixed_ref.hpp"

using namespace oref;

namespace {

constexpr uint32_t MAXNBR = 48;
constexpr int32_t H = 131072;
constexpr int32_t H2Q = 262144;
constexpr uint32_t LUT_SHIFT = 7;
constexpr uint32_t OFF_W = 0, OFF_G = 2049, OFF_P = 4098, OFF_PG = 4611, OFF_BW = 5124, OFF_BG = 6148;
constexpr int32_t BALL_R2MAX = 4194304;
constexpr uint32_t OFF_MAT = 7172, MAT_WORDS = 4;
constexpr int32_t RHO0 = 1048576, C_MAX = 1048576, LAM_MAX = 4194304, EPS_S = 524288;
constexpr unsigned GRAD_S = 16, S_ACC_S = 16, LAM_S = 24, DP_S = 20;
constexpr int32_t S_TERM = 8388608;
constexpr int32_t MARGIN = 32768;

struct V3 { int32_t x, y, z; };
struct Part { int32_t p[4], v[4], pr[4], dp[4]; };  // 64 B (3d/sim3d.wgsl:46-51)
static_assert(sizeof(Part) == 64, "Part");

struct World3 {
  uint32_t maxN = 0, maxCells = 0;
  uint32_t n = 0, gx = 1, gy = 1, gz = 1, cellTotal = 1;
  int iters = 3;
  int32_t stirAmp = 0, stirPhase = 0;
  int32_t ball[3] = {0, 0, 0};
  int32_t ballReach = 0;
  uint32_t ballOn = 0;
  uint32_t substep = 0;
  std::vector<Part> parts;
  std::vector<uint32_t> cellCount, cellStart, cellOf, sortedIds, nbr, nbrN;
  std::vector<int32_t> luts;
  std::atomic<uint32_t> dbg[16];
  uint32_t chain[8], acc[8];
};

inline int32_t world_max(const World3& W, int axis) {
  return axis == 0 ? (int32_t)W.gx * H : (axis == 1 ? (int32_t)W.gy * H : (int32_t)W.gz * H);
}
inline int32_t mat_grav(const World3& W, uint32_t m) { return W.luts[OFF_MAT + m * MAT_WORDS + 0]; }
inline int32_t mat_rhoW(const World3& W, uint32_t m) { return W.luts[OFF_MAT + m * MAT_WORDS + 1]; }
inline int32_t mat_damp(const World3& W, uint32_t m) { return W.luts[OFF_MAT + m * MAT_WORDS + 2]; }
inline V3 clamp_world(const World3& W, V3 p) {
  return V3{clampi(p.x, MARGIN, world_max(W, 0) - MARGIN), clampi(p.y, MARGIN, world_max(W, 1) - MARGIN),
            clampi(p.z, MARGIN, world_max(W, 2) - MARGIN)};
}
inline uint32_t cell_index3(const World3& W, V3 c) {
  return ((uint32_t)c.z * W.gy + (uint32_t)c.y) * W.gx + (uint32_t)c.x;
}
inline V3 cell_coord3(V3 p) { return V3{p.x / H, p.y / H, p.z / H}; }

// wall_of (3d/sim3d.wgsl:129-141)
inline void wall_of(const World3& W, V3 p, uint32_t w, int32_t gm, int32_t& dist, V3& dir) {
  uint32_t axis = w >> 1;
  bool hi = (w & 1u) == 1u;
  int32_t pos = axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
  int32_t d = pos;
  if (hi) d = wsub(world_max(W, (int)axis), pos);
  int32_t s = hi ? gm : fp_neg_sat(gm);
  dir = V3{0, 0, 0};
  if (axis == 0) dir.x = s; else if (axis == 1) dir.y = s; else dir.z = s;
  dist = std::max(d, 0);
}

// ball_term (3d/sim3d.wgsl:152-173)
inline bool ball_term(const World3& W, V3 pi, int32_t& w, V3& g) {
  if (W.ballOn != 1u) return false;
  V3 dc{wsub(W.ball[0], pi.x), wsub(W.ball[1], pi.y), wsub(W.ball[2], pi.z)};
  if (fp_abs(dc.x) >= W.ballReach || fp_abs(dc.y) >= W.ballReach || fp_abs(dc.z) >= W.ballReach) return false;
  V3 q{dc.x >> 8, dc.y >> 8, dc.z >> 8};
  int32_t r2q = wadd(wadd(wmul(q.x, q.x), wmul(q.y, q.y)), wmul(q.z, q.z));
  if (r2q >= BALL_R2MAX) return false;
  uint32_t bi = (uint32_t)r2q >> 12;
  int32_t gt = W.luts[OFF_BG + bi];
  w = W.luts[OFF_BW + bi];
  g = V3{fp_mul_s(gt, dc.x, GRAD_S), fp_mul_s(gt, dc.y, GRAD_S), fp_mul_s(gt, dc.z, GRAD_S)};
  return true;
}

#define PFOR3(N) _Pragma("omp parallel for schedule(static)") for (int64_t ii = 0; ii < (int64_t)(N); ++ii)

inline int32_t s_term(int32_t a) { return clampi(fp_mul_s(a, a, S_ACC_S), 0, S_TERM); }

// The next substep up to and including `stop`: 0 = buildNbr, 1 = iteration
// 0's solveDp (dp not yet applied), 2 = the whole substep.
void substep3_until(World3& W, int stop) {
  g_dbg = W.dbg;
  const uint32_t n = W.n;
  // predict (3d/sim3d.wgsl:188-215)
  PFOR3(n) {
    uint32_t i = (uint32_t)ii;
    Part& P = W.parts[i];
    uint32_t mi = (uint32_t)P.p[3];
    V3 v{P.v[0], P.v[1], P.v[2]};
    v.y = fp_sub_sat(v.y, mat_grav(W, mi));
    if (W.stirAmp != 0) {
      int32_t tri = std::abs(((W.stirPhase + 32768) & 131071) - 65536) - 32768;
      v.x = fp_add_sat(v.x, fp_mul_s(tri, W.stirAmp, 16));
    }
    int32_t dmp = mat_damp(W, mi);
    v = V3{fp_mul_s(v.x, dmp, 16), fp_mul_s(v.y, dmp, 16), fp_mul_s(v.z, dmp, 16)};
    V3 pr = clamp_world(W, V3{fp_add_sat(P.p[0], v.x), fp_add_sat(P.p[1], v.y), fp_add_sat(P.p[2], v.z)});
    P.v[0] = v.x; P.v[1] = v.y; P.v[2] = v.z; P.v[3] = 0;
    P.pr[0] = pr.x; P.pr[1] = pr.y; P.pr[2] = pr.z; P.pr[3] = 0;
  }
  // gridCount + exact scan + stable placement == canonicalize's ascending ids
  for (uint32_t c = 0; c < W.cellTotal; ++c) W.cellCount[c] = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const Part& P = W.parts[i];
    uint32_t c = cell_index3(W, cell_coord3(V3{P.pr[0], P.pr[1], P.pr[2]}));
    W.cellOf[i] = c;
    W.cellCount[c]++;
  }
  uint32_t run = 0;
  for (uint32_t c = 0; c < W.cellTotal; ++c) { W.cellStart[c] = run; run += W.cellCount[c]; W.cellCount[c] = 0; }
  W.cellStart[W.cellTotal] = n;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t c = W.cellOf[i];
    W.sortedIds[W.cellStart[c] + W.cellCount[c]++] = i;
  }
  // buildNbr (3d/sim3d.wgsl:341-390): dz, dy, dx, ascending id
  PFOR3(n) {
    uint32_t i = (uint32_t)ii;
    const Part& P = W.parts[i];
    V3 pi{P.pr[0], P.pr[1], P.pr[2]};
    V3 c = cell_coord3(pi);
    size_t base = (size_t)i * MAXNBR;
    uint32_t cnt = 0, dropped = 0;
    for (int dz = -1; dz <= 1; ++dz) {
      int cz = c.z + dz;
      if (cz < 0 || cz >= (int)W.gz) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        int cy = c.y + dy;
        if (cy < 0 || cy >= (int)W.gy) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          int cx = c.x + dx;
          if (cx < 0 || cx >= (int)W.gx) continue;
          uint32_t cc = cell_index3(W, V3{cx, cy, cz});
          for (uint32_t s = W.cellStart[cc]; s < W.cellStart[cc + 1]; ++s) {
            uint32_t j = W.sortedIds[s];
            if (j == i) continue;
            const Part& Q = W.parts[j];
            V3 d{wsub(Q.pr[0], pi.x), wsub(Q.pr[1], pi.y), wsub(Q.pr[2], pi.z)};
            if (fp_abs(d.x) >= H || fp_abs(d.y) >= H || fp_abs(d.z) >= H) continue;
            V3 q{d.x >> 8, d.y >> 8, d.z >> 8};
            int32_t r2q = wadd(wadd(wmul(q.x, q.x), wmul(q.y, q.y)), wmul(q.z, q.z));
            if (r2q >= H2Q) continue;
            if (cnt >= MAXNBR) { dropped++; continue; }
            W.nbr[base + cnt++] = j;
          }
        }
      }
    }
    if (dropped) W.dbg[6].fetch_add(dropped, std::memory_order_relaxed);
    W.nbrN[i] = cnt;
  }
  if (stop == 0) return;
  for (int it = 0; it < W.iters; ++it) {
    // density (3d/sim3d.wgsl:397-468)
    PFOR3(n) {
      uint32_t i = (uint32_t)ii;
      Part& P = W.parts[i];
      V3 pi{P.pr[0], P.pr[1], P.pr[2]};
      uint32_t cnt = W.nbrN[i];
      size_t base = (size_t)i * MAXNBR;
      int32_t rho = wmul(W.luts[OFF_W], wadd(256, mat_rhoW(W, (uint32_t)P.p[3]))) >> 8;
      int32_t S = 0;
      V3 sg{0, 0, 0};
      for (uint32_t k = 0; k < cnt; ++k) {
        uint32_t j = W.nbr[base + k];
        const Part& Q = W.parts[j];
        V3 d{wsub(Q.pr[0], pi.x), wsub(Q.pr[1], pi.y), wsub(Q.pr[2], pi.z)};
        V3 q{d.x >> 8, d.y >> 8, d.z >> 8};
        int32_t r2q = wadd(wadd(wmul(q.x, q.x), wmul(q.y, q.y)), wmul(q.z, q.z));
        if (r2q >= H2Q) continue;
        uint32_t ui = (uint32_t)r2q >> LUT_SHIFT;
        rho = wadd(rho, wmul(W.luts[OFF_W + ui], wadd(256, mat_rhoW(W, (uint32_t)Q.p[3]))) >> 8);
        int32_t gt = W.luts[OFF_G + ui];
        V3 g{fp_mul_s(gt, d.x, GRAD_S), fp_mul_s(gt, d.y, GRAD_S), fp_mul_s(gt, d.z, GRAD_S)};
        sg = V3{wadd(sg.x, g.x), wadd(sg.y, g.y), wadd(sg.z, g.z)};
        S = wadd(wadd(wadd(S, s_term(g.x)), s_term(g.y)), s_term(g.z));
      }
      for (uint32_t w = 0; w < 6; ++w) {
        int32_t dist;
        V3 dir;
        wall_of(W, pi, w, 0, dist, dir);
        if (dist >= H) continue;
        uint32_t b = (uint32_t)dist >> 8;
        rho = wadd(rho, W.luts[OFF_P + b]);
        V3 gv;
        wall_of(W, pi, w, W.luts[OFF_PG + b], dist, gv);
        sg = V3{wadd(sg.x, gv.x), wadd(sg.y, gv.y), wadd(sg.z, gv.z)};
        S = wadd(wadd(wadd(S, s_term(gv.x)), s_term(gv.y)), s_term(gv.z));
      }
      int32_t bw;
      V3 bg;
      if (ball_term(W, pi, bw, bg)) {
        rho = wadd(rho, bw);
        sg = V3{wadd(sg.x, bg.x), wadd(sg.y, bg.y), wadd(sg.z, bg.z)};
        S = wadd(wadd(wadd(S, s_term(bg.x)), s_term(bg.y)), s_term(bg.z));
      }
      S = wadd(wadd(wadd(S, s_term(sg.x)), s_term(sg.y)), s_term(sg.z));
      int32_t Cq = clampi(fp_shr_rne(wsub(rho, RHO0), 4), 0, C_MAX);
      P.dp[3] = rho;
      P.pr[3] = clampi(fp_neg_sat(fp_divshift(Cq, wadd(S, EPS_S), LAM_S)), fp_neg_sat(LAM_MAX), LAM_MAX);
    }
    // solveDp (3d/sim3d.wgsl:476-527): saturating running sum, canonical order
    PFOR3(n) {
      uint32_t i = (uint32_t)ii;
      Part& P = W.parts[i];
      V3 pi{P.pr[0], P.pr[1], P.pr[2]};
      int32_t lami = P.pr[3];
      uint32_t cnt = W.nbrN[i];
      size_t base = (size_t)i * MAXNBR;
      V3 acc{0, 0, 0};
      for (uint32_t k = 0; k < cnt; ++k) {
        uint32_t j = W.nbr[base + k];
        const Part& Q = W.parts[j];
        V3 d{wsub(Q.pr[0], pi.x), wsub(Q.pr[1], pi.y), wsub(Q.pr[2], pi.z)};
        V3 q{d.x >> 8, d.y >> 8, d.z >> 8};
        int32_t r2q = wadd(wadd(wmul(q.x, q.x), wmul(q.y, q.y)), wmul(q.z, q.z));
        if (r2q >= H2Q) continue;
        uint32_t ui = (uint32_t)r2q >> LUT_SHIFT;
        int32_t gt = W.luts[OFF_G + ui];
        V3 g{fp_mul_s(gt, d.x, GRAD_S), fp_mul_s(gt, d.y, GRAD_S), fp_mul_s(gt, d.z, GRAD_S)};
        int32_t lsum = fp_add_sat(lami, Q.pr[3]);
        acc = V3{fp_add_sat(acc.x, fp_mul_s(lsum, g.x, DP_S)), fp_add_sat(acc.y, fp_mul_s(lsum, g.y, DP_S)),
                 fp_add_sat(acc.z, fp_mul_s(lsum, g.z, DP_S))};
      }
      for (uint32_t w = 0; w < 6; ++w) {
        int32_t dist;
        V3 dir;
        wall_of(W, pi, w, 0, dist, dir);
        if (dist >= H) continue;
        uint32_t b = (uint32_t)dist >> 8;
        V3 gv;
        wall_of(W, pi, w, W.luts[OFF_PG + b], dist, gv);
        acc = V3{fp_add_sat(acc.x, fp_mul_s(lami, gv.x, DP_S)), fp_add_sat(acc.y, fp_mul_s(lami, gv.y, DP_S)),
                 fp_add_sat(acc.z, fp_mul_s(lami, gv.z, DP_S))};
      }
      int32_t bw;
      V3 bg;
      if (ball_term(W, pi, bw, bg)) {
        acc = V3{fp_add_sat(acc.x, fp_mul_s(lami, bg.x, DP_S)), fp_add_sat(acc.y, fp_mul_s(lami, bg.y, DP_S)),
                 fp_add_sat(acc.z, fp_mul_s(lami, bg.z, DP_S))};
      }
      P.dp[0] = acc.x; P.dp[1] = acc.y; P.dp[2] = acc.z;
    }
    if (stop == 1) return;
    // applyDp (3d/sim3d.wgsl:529-537)
    PFOR3(n) {
      Part& P = W.parts[(size_t)ii];
      V3 pr = clamp_world(W, V3{fp_add_sat(P.pr[0], P.dp[0]), fp_add_sat(P.pr[1], P.dp[1]),
                                fp_add_sat(P.pr[2], P.dp[2])});
      P.pr[0] = pr.x; P.pr[1] = pr.y; P.pr[2] = pr.z;
    }
  }
  // integrate (3d/sim3d.wgsl:542-552)
  PFOR3(n) {
    Part& P = W.parts[(size_t)ii];
    P.v[0] = fp_sub_sat(P.pr[0], P.p[0]);
    P.v[1] = fp_sub_sat(P.pr[1], P.p[1]);
    P.v[2] = fp_sub_sat(P.pr[2], P.p[2]);
    P.v[3] = 0;
    P.p[0] = P.pr[0]; P.p[1] = P.pr[1]; P.p[2] = P.pr[2];
  }
  W.substep++;
}

void substep3(World3& W) { substep3_until(W, 2); }

// digestState over 16-word records, hashing words 0..7 (p.xyzw, v.xyzw)
inline uint32_t rotl32(uint32_t x, uint32_t r) { return (x << (r & 31u)) | (x >> ((32u - r) & 31u)); }
inline uint32_t mm_step(uint32_t h, uint32_t k) {
  k *= 0xcc9e2d51u; k = rotl32(k, 15); k *= 0x1b873593u;
  h ^= k; h = rotl32(h, 13);
  return h * 5u + 0xe6546b64u;
}
inline uint32_t fmix32(uint32_t h) {
  h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
  return h;
}
constexpr uint32_t SEED[4] = {0x9e3779b1u, 0x85ebca77u, 0xc2b2ae3du, 0x27d4eb2fu};
constexpr uint32_t SALT[4] = {0x00000000u, 0x51633e2du, 0x9c8f2a17u, 0xd0e5b943u};
constexpr uint32_t CHAIN_INIT[8] = {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u,
                                    0xa4093822u, 0x299f31d0u, 0x082efa98u, 0xec4e6c89u};

void digest3(const World3& W, uint32_t acc[8]) {
  uint32_t a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t i = 0; i < W.n; ++i) {
    const Part& P = W.parts[i];
    uint32_t wds[8] = {(uint32_t)P.p[0], (uint32_t)P.p[1], (uint32_t)P.p[2], (uint32_t)P.p[3],
                       (uint32_t)P.v[0], (uint32_t)P.v[1], (uint32_t)P.v[2], (uint32_t)P.v[3]};
    for (int l = 0; l < 4; ++l) {
      uint32_t h = SEED[l];
      for (int w = 0; w < 8; ++w) h = mm_step(h, wds[w] ^ SALT[l]);
      a[l] += fmix32(h ^ 8u);
      a[4 + l] += fmix32(mm_step(h, i) ^ 9u);
    }
  }
  for (int l = 0; l < 8; ++l) acc[l] = a[l];
}

}  // namespace

extern "C" {

struct or3_world;

or3_world* or3_create(uint32_t max_particles, uint32_t max_cells, const int32_t* luts, uint32_t lut_words,
                      int iters) {
  World3* W = new World3();
  W->maxN = max_particles;
  W->maxCells = max_cells;
  W->iters = iters;
  W->luts.assign(luts, luts + lut_words);
  W->parts.assign(max_particles, Part{});
  W->cellCount.assign(max_cells, 0u);
  W->cellStart.assign((size_t)max_cells + 1, 0u);
  W->cellOf.assign(max_particles, 0u);
  W->sortedIds.assign(max_particles, 0u);
  W->nbr.assign((size_t)max_particles * MAXNBR, 0u);
  W->nbrN.assign(max_particles, 0u);
  for (auto& d : W->dbg) d.store(0);
  return reinterpret_cast<or3_world*>(W);
}

void or3_destroy(or3_world* w) { delete reinterpret_cast<World3*>(w); }

// parts16: n records {p.xyzw (w = material), v.xyzw, pr.xyzw, dp.xyzw}
int or3_load(or3_world* w, uint32_t n, const int32_t* parts16, uint32_t gx, uint32_t gy, uint32_t gz) {
  World3& W = *reinterpret_cast<World3*>(w);
  if (n > W.maxN || (uint64_t)gx * gy * gz > W.maxCells) return -1;
  W.n = n; W.gx = gx; W.gy = gy; W.gz = gz; W.cellTotal = gx * gy * gz;
  W.substep = 0;
  std::memcpy(W.parts.data(), parts16, (size_t)n * 64);
  for (auto& d : W.dbg) d.store(0);
  for (int j = 0; j < 8; ++j) { W.chain[j] = CHAIN_INIT[j]; W.acc[j] = 0; }
  return 0;
}

void or3_set_inputs(or3_world* w, int32_t stirAmp, int32_t stirPhase, uint32_t ballOn, const int32_t* ball3,
                    int32_t ballReach) {
  World3& W = *reinterpret_cast<World3*>(w);
  W.stirAmp = stirAmp;
  W.stirPhase = stirPhase;
  W.ballOn = ballOn;
  if (ball3) { W.ball[0] = ball3[0]; W.ball[1] = ball3[1]; W.ball[2] = ball3[2]; }
  W.ballReach = ballReach;
}

void or3_step(or3_world* w, uint32_t count, uint32_t* acc_stream) {
  World3& W = *reinterpret_cast<World3*>(w);
  for (uint32_t s = 0; s < count; ++s) {
    substep3(W);
    digest3(W, W.acc);
    uint32_t next[8];
    for (int j = 0; j < 8; ++j) next[j] = mm_step(W.chain[j], W.acc[j] ^ rotl32(W.acc[(j + 1) & 7], 7));
    for (int j = 0; j < 8; ++j) W.chain[j] = next[j];
    if (acc_stream)
      for (int l = 0; l < 8; ++l) acc_stream[(size_t)s * 8 + l] = W.acc[l];
  }
}

void or3_digest(or3_world* w, uint32_t acc[8]) { digest3(*reinterpret_cast<World3*>(w), acc); }
void or3_chain(or3_world* w, uint32_t out[8]) {
  World3& W = *reinterpret_cast<World3*>(w);
  for (int j = 0; j < 8; ++j) out[j] = W.chain[j];
}
void or3_read(or3_world* w, int32_t* out16) {
  World3& W = *reinterpret_cast<World3*>(w);
  std::memcpy(out16, W.parts.data(), (size_t)W.n * 64);
}
void or3_debug_counters(or3_world* w, uint32_t out[16]) {
  World3& W = *reinterpret_cast<World3*>(w);
  for (int j = 0; j < 16; ++j) out[j] = W.dbg[j].load();
}

// The 3D stir event (this build's definition; the reference has no 3D
// impulse), applied between substeps on the current state: slots in
// ascending order whose p lies within r of (cx, cy, cz) (IEEE double test,
// as applyImpulse's src/engine.js:6077-6098 in 2D), at most `limit`, get
// (ix, iy, iz) added to v, each component clamped to +-65536.  Returns the
// hit count.
uint32_t or3_apply_impulse(or3_world* w, int32_t cx, int32_t cy, int32_t cz, int32_t r, int32_t ix, int32_t iy,
                           int32_t iz, uint32_t limit) {
  World3& W = *reinterpret_cast<World3*>(w);
  const double r2 = (double)r * (double)r;
  uint32_t hits = 0;
  auto cl = [](int64_t x) { return (int32_t)std::max<int64_t>(-65536, std::min<int64_t>(65536, x)); };
  for (uint32_t i = 0; i < W.n && hits < limit; ++i) {
    Part& P = W.parts[i];
    double dx = (double)P.p[0] - (double)cx, dy = (double)P.p[1] - (double)cy, dz = (double)P.p[2] - (double)cz;
    double q = dx * dx;
    q = q + dy * dy;
    q = q + dz * dz;
    if (q > r2) continue;
    P.v[0] = cl((int64_t)P.v[0] + ix);
    P.v[1] = cl((int64_t)P.v[1] + iy);
    P.v[2] = cl((int64_t)P.v[2] + iz);
    hits++;
  }
  return hits;
}

// Test probe: the next substep through buildNbr only.  Outputs the grid
// (cellStart, cellTotal + 1 words; sortedIds, n words), the lists (48 words
// per slot, list order), their counts and the pairs truncation dropped; the
// predicted positions are in the records (or3_read, words 8..10).  The world
// is left mid-substep (only for tests that discard it).
void or3_grid_probe(or3_world* w, uint32_t* cell_start, uint32_t* sorted_ids, uint32_t* nbr48, uint32_t* nbr_n,
                    uint32_t* dropped) {
  World3& W = *reinterpret_cast<World3*>(w);
  const uint32_t before = W.dbg[6].load();
  substep3_until(W, 0);
  *dropped = W.dbg[6].load() - before;
  std::memcpy(cell_start, W.cellStart.data(), ((size_t)W.cellTotal + 1) * 4);
  std::memcpy(sorted_ids, W.sortedIds.data(), (size_t)W.n * 4);
  std::memcpy(nbr48, W.nbr.data(), (size_t)W.n * MAXNBR * 4);
  std::memcpy(nbr_n, W.nbrN.data(), (size_t)W.n * 4);
}

// Test probe: the next substep through iteration 0's solveDp; the records
// (or3_read) then hold pr, lambda (pr.w), rho (dp.w) and the unapplied dp.
void or3_dp_probe(or3_world* w) { substep3_until(*reinterpret_cast<World3*>(w), 1); }

// The reference's own 3D digest: 256 stripes of FNV-1a over p.xyz, v.xyz in
// slot order (3d/sim3d.wgsl:571-590), folded into 8 lanes on the host
// (3d/app.js:731-742).  Returns the 8 lanes.
void or3_fnv_digest(or3_world* w, uint32_t lanes_out[8]) {
  World3& W = *reinterpret_cast<World3*>(w);
  auto fnv = [](uint32_t h, uint32_t x) {
    h = (h ^ (x & 0xffu)) * 16777619u;
    h = (h ^ ((x >> 8) & 0xffu)) * 16777619u;
    h = (h ^ ((x >> 16) & 0xffu)) * 16777619u;
    h = (h ^ ((x >> 24) & 0xffu)) * 16777619u;
    return h;
  };
  uint32_t words[256];
  uint32_t per = (W.n + 255u) / 256u;
  for (uint32_t k = 0; k < 256; ++k) {
    uint32_t lo = k * per, hi = std::min(lo + per, W.n);
    uint32_t h = 2166136261u;
    for (uint32_t i = lo; i < hi; ++i) {
      const Part& P = W.parts[i];
      h = fnv(h, (uint32_t)P.p[0]); h = fnv(h, (uint32_t)P.p[1]); h = fnv(h, (uint32_t)P.p[2]);
      h = fnv(h, (uint32_t)P.v[0]); h = fnv(h, (uint32_t)P.v[1]); h = fnv(h, (uint32_t)P.v[2]);
    }
    words[k] = h;
  }
  uint32_t lanes[8];
  for (uint32_t L = 0; L < 8; ++L) lanes[L] = 2166136261u ^ (uint32_t)((uint64_t)L * 0x9e3779b9ull);
  for (uint32_t k = 0; k < 256; ++k)
    for (uint32_t L = 0; L < 8; ++L) {
      uint32_t h = lanes[L] ^ words[k];
      h = h * 16777619u;
      h = h ^ (h >> 13);
      lanes[L] = h;
    }
  for (int L = 0; L < 8; ++L) lanes_out[L] = lanes[L];
}

}  // extern "C"

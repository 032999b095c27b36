// ORACLE — test infrastructure only.  A scalar CPU restatement of the
// reference's 2D deterministic PBF substep (src/sim.wgsl) and its state hash
// (src/hash/statehash.wgsl), driven in the reference's pass order
// (src/engine.js:4003-4155).  Only tests/, the smoke check and the CPU-baseline
// leg of bench.py may load it; nothing in the product path links it.
//
// Parallelism: every pass is a per-particle gather whose result is a pure
// function of the pass inputs, exactly as on the GPU, so OpenMP over particles
// gives bit-identical results at any thread count.
//
// Not restated (inert in every supported scene): the rigid-body passes
// (rigidStep/rigidSolve/rigidMember, 0 bodies), jets (nJets = 0), weld/fuse
// gestures (weldHold = fuseTag = 0).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "fixed_ref.hpp"

namespace oref {
std::atomic<uint32_t>* g_dbg = nullptr;
}
using namespace oref;

namespace {

// ---- constants (src/sim.wgsl:55-171, src/engine.js:242-512) -------------
constexpr int32_t H = 163840;
constexpr int32_t RHO0 = 1048576;
constexpr int32_t TWO_RHO0 = 2097152;
constexpr unsigned PRE = 4;
constexpr int32_t HS = 10240;
constexpr int32_t H2S = 104857600;
constexpr unsigned KSH = 15;
constexpr unsigned GF = 22;
constexpr unsigned CF = 22;
constexpr unsigned ACC = 8;
constexpr uint32_t MAXNBR = 48;
constexpr uint32_t TN = 3202;
constexpr uint32_t OFF_W = 0, OFF_G = 3202, OFF_C = 6404, OFF_A = 9606, OFF_S = 12808;
constexpr int32_t LAM_MAX = 4194304, NEG_LAM_MAX = -4194304;
constexpr int32_t S_TERM = 16777216;
constexpr int32_t DP_TERM = 2097152, NEG_DP_TERM = -2097152;
constexpr int32_t XS_TERM = 2097152, NEG_XS_TERM = -2097152;
constexpr int32_t ST_TERM = 4194304, NEG_ST_TERM = -4194304;
constexpr int32_t RINV_MAX = 262144;
constexpr int32_t WALL = 32768;
constexpr int32_t C_MAX = 58000;
constexpr int32_t RHO_FLOOR_RINV = 32768;
constexpr int32_t RHO_FLOOR_K = 65536;
constexpr int32_t AMBIENT = 1310720;
constexpr int32_t TEMP_MIN = -33554432, TEMP_MAX = 536870912;
constexpr int32_t DT_CLAMP = 16777216, NEG_DT_CLAMP = -16777216;
constexpr int32_t HT_TERM = 16777216, NEG_HT_TERM = -16777216;
constexpr int32_t CORR_TERM = 16777216;
constexpr int32_t CORR_NEVER = 2147483647;
constexpr uint32_t FLAG_BURNING = 1, FLAG_DEAD = 2, FLAG_ANCHOR = 4, FLAG_RIGID = 16,
                   FLAG_CONGEAL = 32, FLAG_BURST = 64, FLAG_SEAL = 128;
constexpr uint32_t WELD_TAG_MASK = 0x000fff00u;
constexpr uint32_t NBR_SEAL_BIT = 0x80000000u, NBR_EXT_SEAL_BIT = 0x40000000u,
                   NBR_CONGEAL_BIT = 0x20000000u;
constexpr unsigned FPRE = 5, RPRE = 3, BPRE = 5;
constexpr int32_t D2_FLOOR = 262144;
constexpr int32_t FR_TERM = 16777216, NEG_FR_TERM = -16777216;
constexpr uint32_t BOND_CAP = 48;
constexpr int32_t BOND_R2 = 17623204;
constexpr int32_t HARD_BREAK2 = 26214400;
constexpr int32_t L0_MAX = 5120, L0_MIN = 512;
constexpr uint32_t BOND_J_MASK = 0x0001ffffu, BOND_L0_SH = 17, BOND_L0_MASK = 0x3ffe0000u,
                   BOND_TOMB = 0x40000000u;
constexpr int32_t BD_TERM = 16777216, NEG_BD_TERM = -16777216;
constexpr uint32_t MAT_PHOTO_ROW = 15;
constexpr int32_t SEAL_SEG_MAX2 = 8820900, SEAL_SEG_MIN2 = 670761, SEAL_KEEP4 = 717,
                  SEAL_REACH = 95027, SEAL_STRESS4 = 2560;
constexpr int32_t SOLID_NONE = 0;
constexpr uint32_t MAT_WORDS = 48, MAT_MAX = 64;
constexpr uint32_t BOND_PERIOD = 4;

// Material row word indices (src/sim.wgsl:315-431).
enum MatW {
  M_EPS = 0, M_DPMAX, M_XSPH, M_PHANTOM, M_GCOH, M_GCUR, M_ADH, M_GRAV,
  M_COND, M_COOL, M_HEATCAP, M_MELTPT, M_MELTTO, M_FREEZEPT, M_FREEZETO,
  M_BOILPT, M_BOILTO, M_CONDPT, M_CONDTO, M_IGNITE, M_FUEL0, M_BURNRATE,
  M_BURNHEAT, M_BURNTO, M_SPLITTO, M_SPLITPCT, M_SPAWNT, M_TINT, M_CORRODE,
  M_CORRPT, M_CORRTO, M_FRIC, M_BONDK, M_BONDBREAK, M_BONDREFORM, M_BONDFORMV,
  M_BONDFORMP, M_BONDMU, M_BONDGROUP, M_RAMPLO, M_RAMPHI, M_RAMPINV,
  M_RAMPSIGN, M_SOLIDMODE, M_TEMPTARGET, M_TEMPDRIVE, M_BURSTP, M_SEALMODE
};

struct V2 { int32_t x, y; };

// Particle: 8 words, 32 B (src/sim.wgsl:181-188).
struct Particle {
  V2 pos, vel;
  uint32_t matId, flags;
  int32_t temp, fuel;
};
static_assert(sizeof(Particle) == 32, "Particle layout");

// Derived: 16 words, 64 B (src/sim.wgsl:195-207).
struct Derived {
  V2 pred, dp, vtmp, nrm, stAcc;
  int32_t rho, lam, rinv, heat, corr;
  uint32_t pad1;
};
static_assert(sizeof(Derived) == 64, "Derived layout");

struct World {
  // capacity
  uint32_t maxN = 0, maxCells = 0;
  // params (src/sim.wgsl:232-300)
  uint32_t n = 0, nFluid = 0, gridW = 1, gridH = 1, cellTotal = 1, cellShift = 18;
  int32_t originX = 0, originY = 0, domW = 0, domH = 0, vmax = 65536;
  uint32_t weldHold = 0, fuseTag = 0;
  int iters = 4;
  // engine host state (src/engine.js:3786-3820)
  uint32_t substep = 0;          // `_substep`: substeps encoded since load
  bool bondDirty = true;
  int parity = 0;
  // buffers
  std::vector<Particle> state[2];
  std::vector<Derived> derived;
  std::vector<uint32_t> cellCount, cellStart, cellOf, sortedIds, nbr, nbrN, bond, bondCand;
  std::vector<int32_t> luts;
  std::vector<int32_t> mats;      // MAT_MAX * 48
  std::atomic<uint32_t> dbg[16];
  uint32_t chain[8];
  uint32_t acc[8];

  const int32_t* M(uint32_t m) const { return &mats[(size_t)m * MAT_WORDS]; }
};

// ---- shared helpers (src/sim.wgsl:477-594) ------------------------------
inline V2 cell_coord(const World& W, V2 p) {
  int32_t cx = wsub(p.x, W.originX) >> W.cellShift;
  int32_t cy = wsub(p.y, W.originY) >> W.cellShift;
  return V2{clampi(cx, 0, (int32_t)W.gridW - 1), clampi(cy, 0, (int32_t)W.gridH - 1)};
}
inline uint32_t cell_index(const World& W, V2 c) { return (uint32_t)c.y * W.gridW + (uint32_t)c.x; }
inline uint32_t cell_colour(const World& W, uint32_t i) {
  uint32_t c = W.cellOf[i];
  uint32_t cx = c % W.gridW, cy = c / W.gridW;
  return (cx & 1u) | ((cy & 1u) << 1u);
}
// lut_index (src/sim.wgsl:520-527)
inline int32_t lut_index(V2 d) {
  int32_t ax = wadd(fp_abs(d.x), 8) >> PRE;
  int32_t ay = wadd(fp_abs(d.y), 8) >> PRE;
  if (ax > HS || ay > HS) return -1;
  int32_t r2 = ax * ax + ay * ay;
  if (r2 >= H2S) return -1;
  return r2 >> KSH;
}
inline int32_t grad_q16(int32_t gt, int32_t d) { return fp_mul_s(gt, d, GF + 4u); }
inline bool is_boundary(const World& W, uint32_t j) { return j >= W.nFluid; }
inline uint32_t nbr_count(const World& W, uint32_t i) {
  return W.nbrN[i] & ~(NBR_SEAL_BIT | NBR_EXT_SEAL_BIT | NBR_CONGEAL_BIT);
}
inline V2 pred_delta(const World& W, V2 pi, uint32_t j) {
  V2 pj = W.derived[j].pred;
  return V2{fp_sub_sat(pi.x, pj.x), fp_sub_sat(pi.y, pj.y)};
}
inline V2 wall_clamp(const World& W, V2 p) {
  return V2{clampi(p.x, WALL, W.domW - WALL), clampi(p.y, WALL, W.domH - WALL)};
}
inline int32_t mat_mix(int32_t a, int32_t b) { return wadd(a, b) >> 1; }
inline int32_t sym_k(int32_t rhoi, int32_t rhoj) {
  int32_t den = std::max(RHO_FLOOR_K, fp_add_sat(rhoi, rhoj));
  return clampi(fp_divshift(TWO_RHO0, den, 16), 0, RINV_MAX);
}
inline uint32_t id_hash(uint32_t i) {
  uint32_t x = i + 0x9e3779b9u;
  x = (x ^ (x >> 16)) * 0x21f0aaadu;
  x = (x ^ (x >> 15)) * 0x735a2d97u;
  return x ^ (x >> 15);
}
inline uint32_t pick_product(uint32_t primary, uint32_t splitTo, uint32_t splitPct, uint32_t i) {
  if (splitPct == 0u) return primary;
  if ((id_hash(i) & 255u) < splitPct) return splitTo;
  return primary;
}
inline uint32_t weld_tag(uint32_t f) { return f & WELD_TAG_MASK; }
inline bool rg_same_body(uint32_t fi, uint32_t fj) {
  return (fi & FLAG_RIGID) != 0u && (fj & FLAG_RIGID) != 0u && weld_tag(fi) == weld_tag(fj) &&
         weld_tag(fi) != 0u;
}
inline bool seal_class(const int32_t* m) {
  return m[M_BONDK] != 0 && m[M_BONDREFORM] == 0 && m[M_SOLIDMODE] != SOLID_NONE && m[M_PHANTOM] == 0;
}
inline bool seal_liquid(const int32_t* m) {
  return m[M_BONDK] == 0 && m[M_FRIC] == 0 && m[M_GRAV] > 0 && m[M_PHANTOM] == 0;
}
inline bool fuse_row(const int32_t* m, uint32_t id) {
  return m[M_BONDK] != 0 && m[M_BONDREFORM] == 0 && m[M_BONDGROUP] >= 0 && id != MAT_PHOTO_ROW;
}
// bond_k (src/sim.wgsl:1655-1663)
inline int32_t bond_k(const int32_t* m, int32_t T) {
  if (m[M_BONDK] == 0) return 0;
  if (m[M_RAMPSIGN] == 0) return m[M_BONDK];
  int32_t span = fp_sub_sat(m[M_RAMPHI], m[M_RAMPLO]);
  int32_t dT = wclamp(fp_sub_sat(T, m[M_RAMPLO]), 0, span);   // span < 0 possible for a hand-made row
  int32_t u = clampi(fp_mul_s(dT, m[M_RAMPINV], 16), 0, 65536);
  if (m[M_RAMPSIGN] < 0) u = 65536 - u;
  return fp_mul_s(m[M_BONDK], u, 16);
}
inline uint32_t bond_j(uint32_t w) { return w & BOND_J_MASK; }
inline int32_t bond_l0(uint32_t w) { return (int32_t)((w >> BOND_L0_SH) & 0x1fffu); }
inline uint32_t bond_pack(uint32_t j, int32_t l0) {
  return (((uint32_t)l0 & 0x1fffu) << BOND_L0_SH) | (j & BOND_J_MASK);
}
inline int32_t seal_cross_sign(int32_t ax, int32_t ay, int32_t bx, int32_t by) {
  // sign of the wrapped 64-bit a.x*b.y - a.y*b.x (src/sim.wgsl:1994-1999)
  uint64_t d = (uint64_t)((int64_t)ax * by) - (uint64_t)((int64_t)ay * bx);
  if (d == 0) return 0;
  return (d >> 63) ? -1 : 1;
}

#define PFOR(N) _Pragma("omp parallel for schedule(static)") for (int64_t ii = 0; ii < (int64_t)(N); ++ii)

// ---- DISPATCH predict (src/sim.wgsl:958-1000) ---------------------------
void predict(World& W, const Particle* sin) {
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    const int32_t* M = W.M(sin[i].matId);
    if (M[M_PHANTOM] != 0) { W.derived[i].pred = sin[i].pos; continue; }
    if ((sin[i].flags & FLAG_ANCHOR) != 0u) { W.derived[i].pred = sin[i].pos; continue; }
    if ((sin[i].flags & FLAG_RIGID) != 0u) continue;
    V2 st = W.derived[i].stAcc;
    int32_t vmaxN = fp_neg_sat(W.vmax);
    V2 v = sin[i].vel;
    // jet_accel is (0,0) at nJets == 0 and fp_add_sat(x, 0) == x.
    v.x = fp_add_sat(fp_add_sat(v.x, st.x), 0);
    v.y = fp_add_sat(fp_add_sat(fp_add_sat(v.y, st.y), M[M_GRAV]), 0);
    v.x = clampi(v.x, vmaxN, W.vmax);
    v.y = clampi(v.y, vmaxN, W.vmax);
    V2 pos = sin[i].pos;
    V2 pr{fp_add_sat(pos.x, v.x), fp_add_sat(pos.y, v.y)};
    W.derived[i].pred = wall_clamp(W, pr);
  }
}

// ---- gridCount + scan + scatter + canonicalize (src/sim.wgsl:1009-1168) --
// The reference counts with atomics, scans, scatters in atomic-arrival order
// and then canonicalises every cell into ascending id order.  The result —
// cellStart as the exclusive prefix of the counts (cellStart[cellTotal] = n)
// and sortedIds ascending by (cell, id) — is reproduced here by a serial
// stable counting sort over ascending i.
void build_grid(World& W) {
  std::fill(W.cellCount.begin(), W.cellCount.begin() + W.cellTotal, 0u);
  PFOR(W.n) {
    uint32_t i = (uint32_t)ii;
    W.cellOf[i] = cell_index(W, cell_coord(W, W.derived[i].pred));
  }
  for (uint32_t i = 0; i < W.n; ++i) W.cellCount[W.cellOf[i]]++;
  uint32_t nb = (W.cellTotal + 1023u) / 1024u;
  if (nb > 1024u) fp_flag(7);
  uint32_t run = 0;
  for (uint32_t c = 0; c < W.cellTotal; ++c) { W.cellStart[c] = run; run += W.cellCount[c]; }
  W.cellStart[W.cellTotal] = W.n;
  std::fill(W.cellCount.begin(), W.cellCount.begin() + W.cellTotal, 0u);
  for (uint32_t i = 0; i < W.n; ++i) {
    uint32_t c = W.cellOf[i];
    W.sortedIds[W.cellStart[c] + W.cellCount[c]++] = i;
  }
}

// ---- buildNbr (src/sim.wgsl:1190-1254) -----------------------------------
void build_nbr(World& W, const Particle* sin) {
  PFOR(W.n) {
    uint32_t i = (uint32_t)ii;
    if ((sin[i].flags & FLAG_DEAD) != 0u) { W.nbrN[i] = 0u; continue; }
    V2 pi = W.derived[i].pred;
    V2 cc = cell_coord(W, pi);
    size_t base = (size_t)i * MAXNBR;
    uint32_t cnt = 0, dropped = 0;
    bool sealSeen = false, extSeen = false, congealSeen = false;
    for (int32_t dy = -1; dy <= 1; ++dy) {
      int32_t gy = cc.y + dy;
      if (gy < 0 || gy >= (int32_t)W.gridH) continue;
      for (int32_t dx = -1; dx <= 1; ++dx) {
        int32_t gx = cc.x + dx;
        if (gx < 0 || gx >= (int32_t)W.gridW) continue;
        uint32_t c = (uint32_t)gy * W.gridW + (uint32_t)gx;
        for (uint32_t s = W.cellStart[c]; s < W.cellStart[c + 1]; ++s) {
          uint32_t j = W.sortedIds[s];
          if (j == i) continue;
          if ((sin[j].flags & FLAG_DEAD) != 0u) continue;
          if (lut_index(pred_delta(W, pi, j)) < 0) continue;
          if (j < W.nFluid && (sin[j].flags & FLAG_SEAL) != 0u) {
            sealSeen = true;
            if ((sin[i].flags & FLAG_RIGID) != 0u && (sin[j].flags & FLAG_ANCHOR) != 0u &&
                !rg_same_body(sin[i].flags, sin[j].flags))
              extSeen = true;
          }
          if (j < W.nFluid && (sin[j].flags & FLAG_CONGEAL) != 0u) congealSeen = true;
          if (cnt >= MAXNBR) { dropped++; continue; }
          W.nbr[base + cnt] = j;
          cnt++;
        }
      }
    }
    if (dropped) W.dbg[6].fetch_add(dropped, std::memory_order_relaxed);
    W.nbrN[i] = cnt | (sealSeen ? NBR_SEAL_BIT : 0u) | (extSeen ? NBR_EXT_SEAL_BIT : 0u) |
                (congealSeen ? NBR_CONGEAL_BIT : 0u);
  }
}

// ---- solveA (src/sim.wgsl:1270-1318) -------------------------------------
void solve_a(World& W, const Particle* sin) {
  const int32_t* L = W.luts.data();
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    const int32_t* M = W.M(sin[i].matId);
    V2 pi = W.derived[i].pred;
    uint32_t cnt = nbr_count(W, i);
    size_t base = (size_t)i * MAXNBR;
    int32_t rho = L[OFF_W], S = 0, sgx = 0, sgy = 0;
    for (uint32_t k = 0; k < cnt; ++k) {
      uint32_t j = W.nbr[base + k];
      V2 d = pred_delta(W, pi, j);
      int32_t id = lut_index(d);
      if (id < 0) continue;
      rho = wadd(rho, L[OFF_W + id]);
      int32_t gt = L[OFF_G + id];
      int32_t gx = grad_q16(gt, d.x), gy = grad_q16(gt, d.y);
      sgx = wadd(sgx, gx);
      sgy = wadd(sgy, gy);
      S = wadd(wadd(S, clampi(fp_mul_s(gx, gx, 11), 0, S_TERM)), clampi(fp_mul_s(gy, gy, 11), 0, S_TERM));
    }
    S = wadd(wadd(S, clampi(fp_mul_s(sgx, sgx, 11), 0, S_TERM)), clampi(fp_mul_s(sgy, sgy, 11), 0, S_TERM));
    int32_t Cq = clampi(fp_shr_rne(wsub(rho, RHO0), 4), 0, C_MAX);
    W.derived[i].rho = rho;
    W.derived[i].lam = clampi(fp_neg_sat(fp_divshift(Cq, wadd(S, M[M_EPS]), 25)), NEG_LAM_MAX, LAM_MAX);
  }
}

// ---- solveB_at for one colour (src/sim.wgsl:1337-1371) --------------------
void solve_b(World& W, const Particle* sin, uint32_t colour) {
  const int32_t* L = W.luts.data();
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    if (cell_colour(W, i) != colour) continue;
    V2 pi = W.derived[i].pred;
    int32_t li = W.derived[i].lam;
    uint32_t cnt = nbr_count(W, i);
    size_t base = (size_t)i * MAXNBR;
    uint32_t fi = sin[i].flags;
    bool rigidI = (fi & FLAG_RIGID) != 0u;
    int32_t dpx = 0, dpy = 0;
    for (uint32_t k = 0; k < cnt; ++k) {
      uint32_t j = W.nbr[base + k];
      if (rigidI && !is_boundary(W, j) && rg_same_body(fi, sin[j].flags)) continue;
      V2 d = pred_delta(W, pi, j);
      int32_t id = lut_index(d);
      if (id < 0) continue;
      int32_t lj = is_boundary(W, j) ? li : W.derived[j].lam;
      int32_t co = fp_add_sat(fp_add_sat(li, lj), L[OFF_S + id]);
      int32_t gt = L[OFF_G + id];
      dpx = wadd(dpx, clampi(fp_mul_s(co, grad_q16(gt, d.x), 20 - ACC), NEG_DP_TERM, DP_TERM));
      dpy = wadd(dpy, clampi(fp_mul_s(co, grad_q16(gt, d.y), 20 - ACC), NEG_DP_TERM, DP_TERM));
    }
    W.derived[i].dp = V2{dpx, dpy};
  }
}

// ---- applyDp_at (src/sim.wgsl:1418-1436) ---------------------------------
inline void apply_dp_at(World& W, const Particle* sin, uint32_t i) {
  if ((sin[i].flags & FLAG_RIGID) != 0u) return;
  if ((sin[i].flags & FLAG_ANCHOR) != 0u) return;
  const int32_t* M = W.M(sin[i].matId);
  int32_t dm = M[M_DPMAX], dmN = fp_neg_sat(dm);
  V2 dp = W.derived[i].dp;
  V2 d{clampi(fp_shr_rne(dp.x, ACC), dmN, dm), clampi(fp_shr_rne(dp.y, ACC), dmN, dm)};
  V2 pr = W.derived[i].pred;
  V2 np{fp_add_sat(pr.x, d.x), fp_add_sat(pr.y, d.y)};
  W.derived[i].pred = wall_clamp(W, np);
}
void apply_dp(World& W, const Particle* sin, uint32_t colour) {
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    if (cell_colour(W, i) != colour) continue;
    apply_dp_at(W, sin, i);
  }
}
void apply_fric(World& W, const Particle* sin) {
  PFOR(W.nFluid) { apply_dp_at(W, sin, (uint32_t)ii); }
}

// ---- friction + bond gather (src/sim.wgsl:1673-1934) ----------------------
void friction(World& W, const Particle* sin) {
  const int32_t* L = W.luts.data();
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    const int32_t* Mi = W.M(sin[i].matId);
    int32_t fi = Mi[M_FRIC];
    uint32_t flagsI = sin[i].flags;
    bool rigidI = (flagsI & FLAG_RIGID) != 0u;
    if (fi == 0 && Mi[M_BONDK] == 0) {
      if (!rigidI) W.derived[i].dp = V2{0, 0};
      continue;
    }
    V2 pi = W.derived[i].pred;
    V2 dxi{fp_sub_sat(pi.x, sin[i].pos.x), fp_sub_sat(pi.y, sin[i].pos.y)};
    uint32_t cnt = nbr_count(W, i);
    size_t base = (size_t)i * MAXNBR;
    int32_t fx = 0, fy = 0;
    if (fi != 0) {
      for (uint32_t k = 0; k < cnt; ++k) {
        uint32_t j = W.nbr[base + k];
        if (rigidI && !is_boundary(W, j) && rg_same_body(flagsI, sin[j].flags)) continue;
        V2 d = pred_delta(W, pi, j);
        int32_t id = lut_index(d);
        if (id < 0) continue;
        int32_t d4x = fp_shr_rne(d.x, FPRE), d4y = fp_shr_rne(d.y, FPRE);
        int32_t d2 = wadd(wmul(d4x, d4x), wmul(d4y, d4y));
        if (d2 < D2_FLOOR) continue;
        V2 rel = dxi;
        int32_t mu = fi;
        if (!is_boundary(W, j)) {
          V2 pj = W.derived[j].pred;
          rel = V2{fp_sub_sat(dxi.x, fp_sub_sat(pj.x, sin[j].pos.x)),
                   fp_sub_sat(dxi.y, fp_sub_sat(pj.y, sin[j].pos.y))};
          mu = mat_mix(fi, W.M(sin[j].matId)[M_FRIC]);
        }
        if (mu == 0) continue;
        int32_t r4x = fp_shr_rne(rel.x, RPRE), r4y = fp_shr_rne(rel.y, RPRE);
        int32_t dot = wadd(wmul(d4x, r4x), wmul(d4y, r4y));
        int32_t q = fp_divshift(dot, d2, 16);
        int32_t tx = fp_sub_sat(r4x, fp_mul_s(d4x, q, 16));
        int32_t ty = fp_sub_sat(r4y, fp_mul_s(d4y, q, 16));
        int32_t wq = fp_shr_rne(L[OFF_W + id], 4);
        int32_t sx = fp_mul_s(fp_mul_s(tx, wq, 16 - ACC - RPRE), mu, 16);
        int32_t sy = fp_mul_s(fp_mul_s(ty, wq, 16 - ACC - RPRE), mu, 16);
        fx = wsub(fx, clampi(sx, NEG_FR_TERM, FR_TERM));
        fy = wsub(fy, clampi(sy, NEG_FR_TERM, FR_TERM));
      }
    }
    int32_t ki = bond_k(Mi, sin[i].temp);
    int32_t bx = 0, by = 0;
    if (Mi[M_BONDK] != 0 && !rigidI && (flagsI & FLAG_BURST) == 0u) {
      size_t bb = (size_t)i * BOND_CAP;
      for (uint32_t k = 0; k < BOND_CAP; ++k) {
        uint32_t w = W.bond[bb + k];
        if (w == 0u) break;
        if ((w & BOND_TOMB) != 0u) continue;
        uint32_t j = bond_j(w);
        const int32_t* Mj = W.M(sin[j].matId);
        if ((sin[j].flags & FLAG_BURST) != 0u) { W.bond[bb + k] = w | BOND_TOMB; continue; }
        if (Mj[M_BONDK] == 0) { W.bond[bb + k] = w | BOND_TOMB; continue; }
        int32_t L0 = bond_l0(w);
        int32_t L02 = wmul(L0, L0);
        V2 d = pred_delta(W, pi, j);
        int32_t d4x = fp_shr_rne(d.x, BPRE), d4y = fp_shr_rne(d.y, BPRE);
        if (std::abs(d4x) > L0_MAX || std::abs(d4y) > L0_MAX) { W.bond[bb + k] = w | BOND_TOMB; continue; }
        int32_t d2 = wadd(wmul(d4x, d4x), wmul(d4y, d4y));
        if (d2 >= HARD_BREAK2) { W.bond[bb + k] = w | BOND_TOMB; continue; }
        int32_t mb = std::min(Mi[M_BONDBREAK], Mj[M_BONDBREAK]);
        int32_t mub = std::min(Mi[M_BONDMU], Mj[M_BONDMU]);
        if (mb != 0 && mub != 0) {
          int32_t lp = fp_neg_sat(fp_add_sat(W.derived[i].lam, W.derived[j].lam));
          int32_t Pq = 0;
          if (lp > 0) Pq = fp_shr_rne(lp, 4);
          mb = wadd(wadd(65536, wsub(mb, 65536)), fp_mul_s(Pq, mub, 16));
        }
        if (mb != 0 && d2 > fp_mul_s(L02, mb, 16)) { W.bond[bb + k] = w | BOND_TOMB; continue; }
        if (d2 < D2_FLOOR) continue;
        int32_t kj = bond_k(Mj, sin[j].temp);
        int32_t kp = std::min(ki, kj);
        if (kp == 0) continue;
        int32_t C = fp_sub_sat(d2, L02);
        int32_t q = fp_divshift(C, wmul(d2, 4), 16);
        int32_t qk = fp_mul_s(q, kp, 16);
        int32_t gx = fp_mul_s(d4x, qk, 16 - BPRE - ACC);
        int32_t gy = fp_mul_s(d4y, qk, 16 - BPRE - ACC);
        bx = wsub(bx, clampi(gx, NEG_BD_TERM, BD_TERM));
        by = wsub(by, clampi(gy, NEG_BD_TERM, BD_TERM));
      }
    }
    if (rigidI) {
      V2 d0 = W.derived[i].dp;
      W.derived[i].dp = V2{fp_add_sat(d0.x, fx), fp_add_sat(d0.y, fy)};
    } else {
      W.derived[i].dp = V2{fp_add_sat(fx, bx), fp_add_sat(fy, by)};
    }
  }
}

// ---- contain / containOverflow (src/sim.wgsl:2056-2235) ------------------
inline void seal_keep_push(const World& W, V2 pj, V2& pr) {
  int32_t dx4 = fp_shr_rne(fp_sub_sat(pr.x, pj.x), BPRE);
  int32_t dy4 = fp_shr_rne(fp_sub_sat(pr.y, pj.y), BPRE);
  int32_t d2 = wadd(wmul(dx4, dx4), wmul(dy4, dy4));
  if (d2 > 0 && d2 < SEAL_KEEP4 * SEAL_KEEP4) {
    int32_t d4 = (int32_t)isqrt64((uint64_t)(uint32_t)d2);
    if (d4 > 0) {
      int32_t ux = fp_div_s(dx4, d4, 16), uy = fp_div_s(dy4, d4, 16);
      int32_t amt = (SEAL_KEEP4 - d4) << BPRE;
      pr.x = fp_add_sat(pr.x, fp_mul_s(ux, amt, 16));
      pr.y = fp_add_sat(pr.y, fp_mul_s(uy, amt, 16));
    }
  }
  (void)W;
}
inline void seal_segment_test(V2 pj, V2 pk, V2 pos, V2& pr) {
  int32_t ex4 = fp_shr_rne(fp_sub_sat(pk.x, pj.x), BPRE);
  int32_t ey4 = fp_shr_rne(fp_sub_sat(pk.y, pj.y), BPRE);
  int32_t e2 = wadd(wmul(ex4, ex4), wmul(ey4, ey4));
  if (e2 > SEAL_SEG_MAX2 || e2 < SEAL_SEG_MIN2) return;
  if (std::min(pj.x, pk.x) > std::max(pos.x, pr.x) || std::max(pj.x, pk.x) < std::min(pos.x, pr.x) ||
      std::min(pj.y, pk.y) > std::max(pos.y, pr.y) || std::max(pj.y, pk.y) < std::min(pos.y, pr.y))
    return;
  int32_t ex = fp_sub_sat(pk.x, pj.x), ey = fp_sub_sat(pk.y, pj.y);
  int32_t s0 = seal_cross_sign(ex, ey, fp_sub_sat(pos.x, pj.x), fp_sub_sat(pos.y, pj.y));
  int32_t s1 = seal_cross_sign(ex, ey, fp_sub_sat(pr.x, pj.x), fp_sub_sat(pr.y, pj.y));
  if (s0 == 0 || s0 == s1) return;
  int32_t mx = fp_sub_sat(pr.x, pos.x), my = fp_sub_sat(pr.y, pos.y);
  int32_t t0 = seal_cross_sign(mx, my, fp_sub_sat(pj.x, pos.x), fp_sub_sat(pj.y, pos.y));
  int32_t t1 = seal_cross_sign(mx, my, fp_sub_sat(pk.x, pos.x), fp_sub_sat(pk.y, pos.y));
  if (t0 == t1 && t0 != 0) return;
  pr = pos;
}

void contain(World& W, const Particle* sin) {
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    if ((W.nbrN[i] & NBR_SEAL_BIT) == 0u) continue;
    if (nbr_count(W, i) >= MAXNBR) continue;
    const int32_t* Mi = W.M(sin[i].matId);
    if (seal_class(Mi) || Mi[M_PHANTOM] != 0) continue;
    if ((sin[i].flags & (FLAG_ANCHOR | FLAG_RIGID)) != 0u) continue;
    V2 pos = sin[i].pos;
    V2 pr = W.derived[i].pred;
    bool liquidI = seal_liquid(Mi);
    uint32_t cnt = nbr_count(W, i);
    size_t base = (size_t)i * MAXNBR;
    bool touched = false;
    for (uint32_t k = 0; k < cnt; ++k) {
      uint32_t j = W.nbr[base + k];
      if (is_boundary(W, j) || !seal_class(W.M(sin[j].matId))) continue;
      touched = true;
      seal_keep_push(W, W.derived[j].pred, pr);
    }
    if (!touched) { W.derived[i].pred = pr; continue; }
    for (uint32_t k = 0; k < cnt; ++k) {
      uint32_t j = W.nbr[base + k];
      if (is_boundary(W, j)) continue;
      const int32_t* Mj = W.M(sin[j].matId);
      if (!seal_class(Mj)) continue;
      V2 pj = W.derived[j].pred;
      for (uint32_t q = k + 1; q < cnt; ++q) {
        uint32_t k2 = W.nbr[base + q];
        if (is_boundary(W, k2)) continue;
        const int32_t* Mk = W.M(sin[k2].matId);
        if (!seal_class(Mk)) continue;
        if (liquidI && Mj[M_SEALMODE] != 0 && Mk[M_SEALMODE] != 0) continue;
        seal_segment_test(pj, W.derived[k2].pred, pos, pr);
      }
    }
    W.derived[i].pred = pr;
  }
}

void contain_overflow(World& W, const Particle* sin) {
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    if ((W.nbrN[i] & NBR_SEAL_BIT) == 0u) continue;
    if (nbr_count(W, i) < MAXNBR) continue;
    const int32_t* Mi = W.M(sin[i].matId);
    if (seal_class(Mi) || Mi[M_PHANTOM] != 0) continue;
    if ((sin[i].flags & (FLAG_ANCHOR | FLAG_RIGID)) != 0u) continue;
    V2 pos = sin[i].pos;
    V2 pr = W.derived[i].pred;
    bool liquidI = seal_liquid(Mi);
    uint32_t seals[48];
    uint32_t sealN = 0;
    int32_t loX = fp_sub_sat(std::min(pos.x, pr.x), SEAL_REACH);
    int32_t hiX = fp_add_sat(std::max(pos.x, pr.x), SEAL_REACH);
    int32_t loY = fp_sub_sat(std::min(pos.y, pr.y), SEAL_REACH);
    int32_t hiY = fp_add_sat(std::max(pos.y, pr.y), SEAL_REACH);
    V2 cc = cell_coord(W, pr);
    for (int32_t gy = cc.y - 1; gy <= cc.y + 1; ++gy) {
      if (gy < 0 || gy >= (int32_t)W.gridH) continue;
      for (int32_t gx = cc.x - 1; gx <= cc.x + 1; ++gx) {
        if (gx < 0 || gx >= (int32_t)W.gridW) continue;
        uint32_t c = (uint32_t)gy * W.gridW + (uint32_t)gx;
        for (uint32_t t = W.cellStart[c]; t < W.cellStart[c + 1]; ++t) {
          uint32_t j = W.sortedIds[t];
          if (is_boundary(W, j) || !seal_class(W.M(sin[j].matId))) continue;
          V2 pj = W.derived[j].pred;
          if (pj.x < loX || pj.x > hiX || pj.y < loY || pj.y > hiY) continue;
          if (sealN < 48u) seals[sealN++] = j;
        }
      }
    }
    for (uint32_t k = 0; k < sealN; ++k) seal_keep_push(W, W.derived[seals[k]].pred, pr);
    for (uint32_t k = 0; k < sealN; ++k) {
      uint32_t j = seals[k];
      const int32_t* Mj = W.M(sin[j].matId);
      V2 pj = W.derived[j].pred;
      for (uint32_t q = k + 1; q < sealN; ++q) {
        uint32_t k2 = seals[q];
        const int32_t* Mk = W.M(sin[k2].matId);
        if (liquidI && Mj[M_SEALMODE] != 0 && Mk[M_SEALMODE] != 0) continue;
        seal_segment_test(pj, W.derived[k2].pred, pos, pr);
      }
    }
    W.derived[i].pred = pr;
  }
}

// ---- finalize (src/sim.wgsl:2245-2265) -----------------------------------
void finalize(World& W, const Particle* sin, Particle* sout) {
  PFOR(W.n) {
    uint32_t i = (uint32_t)ii;
    Particle outP = sin[i];
    if (i < W.nFluid) {
      int32_t vmaxN = fp_neg_sat(W.vmax);
      V2 pr = W.derived[i].pred;
      V2 v{fp_sub_sat(pr.x, outP.pos.x), fp_sub_sat(pr.y, outP.pos.y)};
      v.x = clampi(v.x, vmaxN, W.vmax);
      v.y = clampi(v.y, vmaxN, W.vmax);
      outP.pos = pr;
      outP.vel = v;
      W.derived[i].vtmp = v;
      W.derived[i].rinv =
          clampi(fp_divshift(RHO0, std::max(W.derived[i].rho, RHO_FLOOR_RINV), 16), 0, RINV_MAX);
    } else {
      outP.vel = V2{0, 0};
    }
    sout[i] = outP;
  }
}

// ---- xsph (src/sim.wgsl:2300-2348) ---------------------------------------
void xsph(World& W, Particle* sout) {
  const int32_t* L = W.luts.data();
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    if ((sout[i].flags & FLAG_RIGID) != 0u) continue;
    if ((sout[i].flags & FLAG_ANCHOR) != 0u) continue;
    int32_t ci = W.M(sout[i].matId)[M_XSPH];
    V2 pi = W.derived[i].pred, vi = W.derived[i].vtmp;
    int32_t rhoi = W.derived[i].rho;
    uint32_t cnt = nbr_count(W, i);
    size_t base = (size_t)i * MAXNBR;
    int32_t ax = 0, ay = 0;
    for (uint32_t k = 0; k < cnt; ++k) {
      uint32_t j = W.nbr[base + k];
      if (is_boundary(W, j)) continue;
      V2 d = pred_delta(W, pi, j);
      int32_t id = lut_index(d);
      if (id < 0) continue;
      int32_t K = sym_k(rhoi, W.derived[j].rho);
      int32_t w = fp_mul_s(L[OFF_W + id], K, 16);
      int32_t cij = mat_mix(ci, W.M(sout[j].matId)[M_XSPH]);
      V2 vj = W.derived[j].vtmp;
      ax = wadd(ax, clampi(fp_mul_s(fp_mul_s(fp_sub_sat(vj.x, vi.x), w, 20 - ACC), cij, 16), NEG_XS_TERM, XS_TERM));
      ay = wadd(ay, clampi(fp_mul_s(fp_mul_s(fp_sub_sat(vj.y, vi.y), w, 20 - ACC), cij, 16), NEG_XS_TERM, XS_TERM));
    }
    V2 v = sout[i].vel;
    sout[i].vel = V2{fp_add_sat(v.x, fp_shr_rne(ax, ACC)), fp_add_sat(v.y, fp_shr_rne(ay, ACC))};
  }
}

// ---- normals (src/sim.wgsl:2357-2384) ------------------------------------
void normals(World& W) {
  const int32_t* L = W.luts.data();
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    V2 pi = W.derived[i].pred;
    uint32_t cnt = nbr_count(W, i);
    size_t base = (size_t)i * MAXNBR;
    int32_t sx = 0, sy = 0;
    for (uint32_t k = 0; k < cnt; ++k) {
      uint32_t j = W.nbr[base + k];
      if (is_boundary(W, j)) continue;
      V2 d = pred_delta(W, pi, j);
      int32_t id = lut_index(d);
      if (id < 0) continue;
      int32_t gt = L[OFF_G + id];
      int32_t rj = W.derived[j].rinv;
      sx = wadd(sx, fp_mul_s(grad_q16(gt, d.x), rj, 16));
      sy = wadd(sy, fp_mul_s(grad_q16(gt, d.y), rj, 16));
    }
    W.derived[i].nrm = V2{fp_mul_s(sx, H, 16), fp_mul_s(sy, H, 16)};
  }
}

// ---- surfaceTension (src/sim.wgsl:2409-2494) -----------------------------
void surface_tension(World& W, const Particle* sout) {
  const int32_t* L = W.luts.data();
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    const int32_t* M = W.M(sout[i].matId);
    V2 pi = W.derived[i].pred, ni = W.derived[i].nrm;
    int32_t rhoi = W.derived[i].rho;
    uint32_t cnt = nbr_count(W, i);
    size_t base = (size_t)i * MAXNBR;
    bool stRigidI = (sout[i].flags & FLAG_RIGID) != 0u;
    int32_t ax = 0, ay = 0;
    for (uint32_t k = 0; k < cnt; ++k) {
      uint32_t j = W.nbr[base + k];
      if (stRigidI && !is_boundary(W, j) && rg_same_body(sout[i].flags, sout[j].flags)) continue;
      V2 d = pred_delta(W, pi, j);
      int32_t id = lut_index(d);
      if (id < 0) continue;
      if (is_boundary(W, j)) {
        int32_t at = L[OFF_A + id];
        ax = wsub(ax, clampi(fp_mul_s(fp_mul_s(at, d.x, CF), M[M_ADH], 16 - ACC), NEG_ST_TERM, ST_TERM));
        ay = wsub(ay, clampi(fp_mul_s(fp_mul_s(at, d.y, CF), M[M_ADH], 16 - ACC), NEG_ST_TERM, ST_TERM));
        continue;
      }
      int32_t K = sym_k(rhoi, W.derived[j].rho);
      int32_t ct = L[OFF_C + id];
      const int32_t* Mj = W.M(sout[j].matId);
      int32_t gCoh = mat_mix(M[M_GCOH], Mj[M_GCOH]);
      int32_t gCur = mat_mix(M[M_GCUR], Mj[M_GCUR]);
      int32_t fx = fp_neg_sat(fp_mul_s(fp_mul_s(ct, d.x, CF), gCoh, 16));
      int32_t fy = fp_neg_sat(fp_mul_s(fp_mul_s(ct, d.y, CF), gCoh, 16));
      V2 nj = W.derived[j].nrm;
      int32_t cx = fp_mul_s(fp_sub_sat(ni.x, nj.x), gCur, 16);
      int32_t cy = fp_mul_s(fp_sub_sat(ni.y, nj.y), gCur, 16);
      // CUR_SIGN < 0: select(cx, fp_neg_sat(cx), true) == fp_neg_sat(cx)
      fx = fp_add_sat(fx, fp_neg_sat(cx));
      fy = fp_add_sat(fy, fp_neg_sat(cy));
      ax = wadd(ax, clampi(fp_mul_s(fx, K, 16 - ACC), NEG_ST_TERM, ST_TERM));
      ay = wadd(ay, clampi(fp_mul_s(fy, K, 16 - ACC), NEG_ST_TERM, ST_TERM));
    }
    W.derived[i].stAcc = V2{fp_shr_rne(ax, ACC), fp_shr_rne(ay, ACC)};
  }
}

// ---- conduct (src/sim.wgsl:2528-2582) ------------------------------------
void conduct(World& W, const Particle* sout) {
  const int32_t* L = W.luts.data();
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    uint32_t mi = sout[i].matId;
    const int32_t* Mi = W.M(mi);
    int32_t ki = Mi[M_COND], Ti = sout[i].temp;
    V2 pi = W.derived[i].pred;
    uint32_t cnt = nbr_count(W, i);
    size_t base = (size_t)i * MAXNBR;
    int32_t q = 0, c = 0;
    for (uint32_t k = 0; k < cnt; ++k) {
      uint32_t j = W.nbr[base + k];
      if (is_boundary(W, j)) continue;
      V2 d = pred_delta(W, pi, j);
      int32_t id = lut_index(d);
      if (id < 0) continue;
      int32_t Wv = L[OFF_W + id];
      uint32_t mj = sout[j].matId;
      const int32_t* Mj = W.M(mj);
      int32_t cij = mat_mix(ki, Mj[M_COND]);
      int32_t wk = fp_mul_s(Wv, cij, 16);
      int32_t dT = clampi(fp_sub_sat(sout[j].temp, Ti), NEG_DT_CLAMP, DT_CLAMP);
      q = wadd(q, clampi(fp_mul_s(dT, wk, 20 - ACC), NEG_HT_TERM, HT_TERM));
      if (mj != mi) {
        int32_t attack = Mj[M_CORRODE];
        int32_t spend = (Mj[M_CORRPT] != CORR_NEVER) ? Mi[M_CORRODE] : 0;
        int32_t str = wadd(attack, spend);
        if (str != 0) c = wadd(c, clampi(fp_mul_s(Wv, str, 20 - ACC), 0, CORR_TERM));
      }
    }
    W.derived[i].heat = q;
    W.derived[i].corr = c;
  }
}

// ---- thermal (src/sim.wgsl:2612-2707) ------------------------------------
void thermal(World& W, Particle* sout) {
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    uint32_t mat0 = sout[i].matId;
    const int32_t* M = W.M(mat0);
    int32_t T = fp_add_sat(sout[i].temp, fp_shr_rne(W.derived[i].heat, ACC));
    T = fp_add_sat(T, fp_mul_s(fp_sub_sat(AMBIENT, T), M[M_COOL], 16));
    int32_t fuel = sout[i].fuel;
    bool burning = (sout[i].flags & FLAG_BURNING) != 0u;
    if (!burning && fuel > 0 && T >= M[M_IGNITE]) burning = true;
    if (burning) {
      T = fp_add_sat(T, M[M_BURNHEAT]);
      fuel = std::max(0, fp_sub_sat(fuel, M[M_BURNRATE]));
    }
    T = clampi(T, TEMP_MIN, TEMP_MAX);
    T = fp_add_sat(T, fp_mul_s(fp_sub_sat(M[M_TEMPTARGET], T), M[M_TEMPDRIVE], 16));
    int32_t dose = fp_shr_rne(W.derived[i].corr, ACC);
    uint32_t mat = mat0;
    bool froze = false;
    uint32_t splitTo = (uint32_t)M[M_SPLITTO], splitPct = (uint32_t)M[M_SPLITPCT];
    if (burning && fuel <= 0) mat = pick_product((uint32_t)M[M_BURNTO], splitTo, splitPct, i);
    else if (dose >= M[M_CORRPT]) mat = pick_product((uint32_t)M[M_CORRTO], splitTo, splitPct, i);
    else if (T >= M[M_BOILPT]) mat = pick_product((uint32_t)M[M_BOILTO], splitTo, splitPct, i);
    else if (T >= M[M_MELTPT]) mat = pick_product((uint32_t)M[M_MELTTO], splitTo, splitPct, i);
    else if (T <= M[M_CONDPT]) mat = pick_product((uint32_t)M[M_CONDTO], splitTo, splitPct, i);
    else if (T <= M[M_FREEZEPT]) {
      mat = pick_product((uint32_t)M[M_FREEZETO], splitTo, splitPct, i);
      froze = true;
    }
    if (mat != mat0) {
      fuel = W.M(mat)[M_FUEL0];
      burning = false;
    }
    uint32_t fl = sout[i].flags & ~(FLAG_BURNING | FLAG_DEAD | FLAG_SEAL);
    if (mat != mat0) fl &= ~(FLAG_ANCHOR | FLAG_CONGEAL | FLAG_BURST);
    if (froze && mat != mat0 && fuse_row(W.M(mat), mat)) fl |= FLAG_CONGEAL;
    if (burning) fl |= FLAG_BURNING;
    if (W.M(mat)[M_PHANTOM] != 0) fl |= FLAG_DEAD;
    if (seal_class(W.M(mat))) fl |= FLAG_SEAL;
    sout[i].matId = mat;
    sout[i].flags = fl;
    sout[i].temp = T;
    sout[i].fuel = fuel;
  }
}

// ---- markCongealNbr (src/sim.wgsl:2714-2730) ------------------------------
void mark_congeal_nbr(World& W, const Particle* sout) {
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    uint32_t cnt = nbr_count(W, i);
    size_t base = (size_t)i * MAXNBR;
    bool seen = false;
    for (uint32_t k = 0; k < cnt; ++k) {
      uint32_t j = W.nbr[base + k];
      if (j < W.nFluid && (sout[j].flags & FLAG_CONGEAL) != 0u) { seen = true; break; }
    }
    W.nbrN[i] = (W.nbrN[i] & ~NBR_CONGEAL_BIT) | (seen ? NBR_CONGEAL_BIT : 0u);
  }
}

// ---- containStress (src/sim.wgsl:2742-2771) -------------------------------
void contain_stress(World& W, Particle* sout) {
  // Reads other particles' flags? No: only matId/pos of j, writes flags of i.
  std::vector<uint32_t> newFlags;
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    const int32_t* Mi = W.M(sout[i].matId);
    if (!seal_class(Mi) || Mi[M_BURSTP] <= 0) continue;
    uint32_t fl = sout[i].flags;
    if ((fl & (FLAG_ANCHOR | FLAG_RIGID)) == 0u) continue;
    V2 pi = sout[i].pos;
    uint32_t cnt = nbr_count(W, i);
    size_t base = (size_t)i * MAXNBR;
    int32_t stress = 0;
    for (uint32_t k = 0; k < cnt; ++k) {
      uint32_t j = W.nbr[base + k];
      if (is_boundary(W, j)) continue;
      const int32_t* Mj = W.M(sout[j].matId);
      if (seal_class(Mj) || Mj[M_PHANTOM] != 0) continue;
      int32_t dx4 = fp_shr_rne(fp_sub_sat(pi.x, sout[j].pos.x), BPRE);
      int32_t dy4 = fp_shr_rne(fp_sub_sat(pi.y, sout[j].pos.y), BPRE);
      int32_t d2 = wadd(wmul(dx4, dx4), wmul(dy4, dy4));
      if (d2 >= SEAL_STRESS4 * SEAL_STRESS4) continue;
      int32_t d4 = (int32_t)isqrt64((uint64_t)(uint32_t)d2);
      stress = wadd(stress, SEAL_STRESS4 - d4);
    }
    if (wmul(stress, 1 << BPRE) > Mi[M_BURSTP]) {
      sout[i].flags = (fl & ~(FLAG_ANCHOR | FLAG_RIGID | WELD_TAG_MASK)) | FLAG_BURST;
    }
  }
}

// ---- bondForm1 (src/sim.wgsl:2794-3027) ----------------------------------
void bond_form1(World& W, const Particle* sout) {
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    size_t cb = (size_t)i * BOND_CAP;
    if ((sout[i].flags & FLAG_RIGID) != 0u) {
      for (uint32_t k = 0; k < BOND_CAP; ++k) W.bondCand[cb + k] = 0u;
      continue;
    }
    bool anchoredI = (sout[i].flags & FLAG_ANCHOR) != 0u;
    const int32_t* Mi = W.M(sout[i].matId);
    uint32_t cn = 0;
    if (Mi[M_BONDK] != 0 && Mi[M_BONDGROUP] >= 0) {
      for (uint32_t k = 0; k < BOND_CAP; ++k) {
        uint32_t w = W.bond[cb + k];
        if (w == 0u) break;
        if ((w & BOND_TOMB) != 0u) continue;
        if (cn >= BOND_CAP) { fp_flag(8); break; }
        W.bondCand[cb + cn] = w;
        cn++;
      }
      uint32_t tagI = weld_tag(sout[i].flags);
      int32_t ki = bond_k(Mi, sout[i].temp);
      bool canReform = (!anchoredI && Mi[M_BONDREFORM] != 0 && ki != 0);
      bool fuseArmed = (W.fuseTag != 0u && fuse_row(Mi, sout[i].matId));
      bool congealI = ((sout[i].flags & FLAG_CONGEAL) != 0u && fuse_row(Mi, sout[i].matId));
      bool congealNear = ((W.nbrN[i] & NBR_CONGEAL_BIT) != 0u && fuse_row(Mi, sout[i].matId));
      if (canReform || (!anchoredI && tagI != 0u) || fuseArmed || congealI || congealNear) {
        uint32_t cnt = nbr_count(W, i);
        size_t nb = (size_t)i * MAXNBR;
        for (uint32_t k = 0; k < cnt; ++k) {
          if (cn >= BOND_CAP) { fp_flag(8); break; }
          uint32_t j = W.nbr[nb + k];
          if (j >= W.nFluid) continue;
          const int32_t* Mj = W.M(sout[j].matId);
          uint32_t fJ = sout[j].flags;
          uint32_t tagJ = weld_tag(fJ);
          bool weld = (!anchoredI && tagI != 0u && tagI == tagJ);
          bool fuse = false;
          if (fuseArmed && (fJ & FLAG_RIGID) == 0u && fuse_row(Mj, sout[j].matId) &&
              (!anchoredI || (fJ & FLAG_ANCHOR) == 0u)) {
            fuse = ((tagI == W.fuseTag && tagJ == 0u) || (tagJ == W.fuseTag && tagI == 0u) ||
                    (tagI == W.fuseTag && tagJ == W.fuseTag));
          }
          bool congealJ = ((fJ & FLAG_CONGEAL) != 0u && fuse_row(Mj, sout[j].matId));
          bool congeal = ((congealI || congealJ) && !weld && !fuse &&
                          sout[j].matId == sout[i].matId && (fJ & FLAG_RIGID) == 0u);
          if (!fuse && !congeal && Mj[M_BONDGROUP] != Mi[M_BONDGROUP]) continue;
          if (!weld && !fuse && !congeal) {
            if (!canReform) continue;
            if (bond_k(Mj, sout[j].temp) == 0) continue;
          }
          bool dup = false;
          for (uint32_t q = 0; q < cn; ++q) {
            if (bond_j(W.bondCand[cb + q]) == j) { dup = true; break; }
          }
          if (dup) continue;
          V2 d{fp_sub_sat(sout[i].pos.x, sout[j].pos.x), fp_sub_sat(sout[i].pos.y, sout[j].pos.y)};
          int32_t d4x = fp_shr_rne(d.x, BPRE), d4y = fp_shr_rne(d.y, BPRE);
          if (std::abs(d4x) > L0_MAX || std::abs(d4y) > L0_MAX) continue;
          int32_t d2 = wadd(wmul(d4x, d4x), wmul(d4y, d4y));
          if (d2 < D2_FLOOR || d2 > BOND_R2) continue;
          if (!weld && !fuse && !congeal) {
            int32_t vl = std::min(Mi[M_BONDFORMV], Mj[M_BONDFORMV]);
            int32_t vl4 = fp_shr_rne(vl, RPRE);
            int32_t r4x = fp_shr_rne(fp_sub_sat(sout[i].vel.x, sout[j].vel.x), RPRE);
            int32_t r4y = fp_shr_rne(fp_sub_sat(sout[i].vel.y, sout[j].vel.y), RPRE);
            if (std::abs(r4x) > vl4 || std::abs(r4y) > vl4) continue;
            if (wadd(wmul(r4x, r4x), wmul(r4y, r4y)) > wmul(vl4, vl4)) continue;
            int32_t pl = std::min(Mi[M_BONDFORMP], Mj[M_BONDFORMP]);
            if (pl != 0 && (fp_neg_sat(W.derived[i].lam) < pl || fp_neg_sat(W.derived[j].lam) < pl)) continue;
          }
          int32_t L0 = (int32_t)isqrt64((uint64_t)(uint32_t)d2);
          if (L0 < L0_MIN || L0 > L0_MAX) continue;
          W.bondCand[cb + cn] = bond_pack(j, L0);
          cn++;
        }
      }
    }
    for (uint32_t k = cn; k < BOND_CAP; ++k) W.bondCand[cb + k] = 0u;
  }
}

// ---- bondForm2 (src/sim.wgsl:3039-3097) ----------------------------------
void bond_form2(World& W, Particle* sout) {
  PFOR(W.nFluid) {
    uint32_t i = (uint32_t)ii;
    size_t cb = (size_t)i * BOND_CAP;
    uint32_t n = 0;
    for (uint32_t k = 0; k < BOND_CAP; ++k) {
      uint32_t w = W.bondCand[cb + k];
      if (w == 0u) break;
      uint32_t j = bond_j(w);
      size_t jb = (size_t)j * BOND_CAP;
      bool ok = false;
      for (uint32_t q = 0; q < BOND_CAP; ++q) {
        uint32_t v = W.bondCand[jb + q];
        if (v == 0u) break;
        if (bond_j(v) == i) { ok = ((v & BOND_L0_MASK) == (w & BOND_L0_MASK)); break; }
      }
      if (ok) { W.bond[cb + n] = w & ~BOND_TOMB; n++; }
    }
    for (uint32_t k = n; k < BOND_CAP; ++k) W.bond[cb + k] = 0u;
    uint32_t fOut = sout[i].flags & ~FLAG_CONGEAL;
    if ((fOut & FLAG_RIGID) == 0u && weld_tag(fOut) != W.weldHold) fOut &= ~WELD_TAG_MASK;
    sout[i].flags = fOut;
  }
}

// ---- state hash (src/hash/statehash.wgsl:54-170) -------------------------
constexpr uint32_t LANE_SEED[4] = {0x9e3779b1u, 0x85ebca77u, 0xc2b2ae3du, 0x27d4eb2fu};
constexpr uint32_t LANE_SALT[4] = {0x00000000u, 0x51633e2du, 0x9c8f2a17u, 0xd0e5b943u};
constexpr uint32_t BOND_SALT = 0x5bd1e995u, BOND_SALT2 = 0x2f5b3c17u;
constexpr uint32_t CHAIN_INIT[8] = {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u,
                                    0xa4093822u, 0x299f31d0u, 0x082efa98u, 0xec4e6c89u};
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

void digest(const World& W, const Particle* st, uint32_t acc[8]) {
  uint32_t a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t n = 8u, m = 9u;
  for (uint32_t i = 0; i < W.n; ++i) {
    const uint32_t* wds = reinterpret_cast<const uint32_t*>(&st[i]);
    uint32_t h[4] = {LANE_SEED[0], LANE_SEED[1], LANE_SEED[2], LANE_SEED[3]};
    for (uint32_t w = 0; w < 8; ++w)
      for (int l = 0; l < 4; ++l) h[l] = mm_step(h[l], wds[w] ^ LANE_SALT[l]);
    uint32_t s[4] = {h[0], h[1], h[2], h[3]}, lz[4] = {h[0], h[1], h[2], h[3]};
    if (i < W.nFluid) {
      uint32_t bn = 0, com = 0, scom = 0;
      for (uint32_t k = 0; k < BOND_CAP; ++k) {
        uint32_t w = W.bond[(size_t)i * BOND_CAP + k];
        if (w == 0u) break;
        com += fmix32(((w >> 17) & 0x1fffu) ^ BOND_SALT);
        scom += fmix32(w ^ BOND_SALT2);
        bn++;
      }
      if (bn != 0u)
        for (int l = 0; l < 4; ++l) { s[l] = mm_step(s[l], bn) ^ com; lz[l] = mm_step(lz[l], bn) ^ scom; }
    }
    for (int l = 0; l < 4; ++l) {
      a[l] += fmix32(s[l] ^ n);
      a[4 + l] += fmix32(mm_step(lz[l], i) ^ m);
    }
  }
  for (int l = 0; l < 8; ++l) acc[l] = a[l];
}

void fold_chain(uint32_t chain[8], const uint32_t acc[8]) {
  uint32_t next[8];
  for (int j = 0; j < 8; ++j) next[j] = mm_step(chain[j], acc[j] ^ rotl32(acc[(j + 1) & 7], 7));
  for (int j = 0; j < 8; ++j) chain[j] = next[j];
}

// ---- one substep in the reference's pass order (src/engine.js:4003-4155) --
void substep(World& W, uint32_t* acc_out) {
  g_dbg = W.dbg;
  const Particle* sin = W.state[W.parity].data();
  Particle* sout = W.state[1 - W.parity].data();
  predict(W, sin);
  build_grid(W);
  build_nbr(W, sin);
  for (int it = 0; it < W.iters; ++it) {
    solve_a(W, sin);
    for (uint32_t c = 0; c < 4; ++c) {
      solve_b(W, sin, c);
      apply_dp(W, sin, c);
    }
    friction(W, sin);
    apply_fric(W, sin);
  }
  contain(W, sin);
  contain_overflow(W, sin);
  finalize(W, sin, sout);
  xsph(W, sout);
  normals(W);
  surface_tension(W, sout);
  conduct(W, sout);
  thermal(W, sout);
  contain_stress(W, sout);
  if ((W.substep % BOND_PERIOD) == 0u || W.bondDirty) {
    mark_congeal_nbr(W, sout);
    bond_form1(W, sout);
    bond_form2(W, sout);
    W.bondDirty = false;
  }
  W.substep++;
  digest(W, sout, W.acc);
  fold_chain(W.chain, W.acc);
  if (acc_out) for (int l = 0; l < 8; ++l) acc_out[l] = W.acc[l];
  W.parity = 1 - W.parity;
}

}  // namespace

// ===================================================================== C API
extern "C" {

struct or_world;

or_world* or_create(uint32_t max_particles, uint32_t max_cells, const int32_t* luts,
                    const int32_t* mats, int iters) {
  World* W = new World();
  W->maxN = max_particles;
  W->maxCells = max_cells;
  W->iters = iters;
  W->luts.assign(luts, luts + 5 * TN);
  W->mats.assign(mats, mats + MAT_MAX * MAT_WORDS);
  for (auto& d : W->dbg) d.store(0);
  W->state[0].assign(max_particles, Particle{});
  W->state[1].assign(max_particles, Particle{});
  W->derived.assign(max_particles, Derived{});
  W->cellCount.assign(max_cells, 0u);
  W->cellStart.assign(max_cells + 1, 0u);
  W->cellOf.assign(max_particles, 0u);
  W->sortedIds.assign(max_particles, 0u);
  W->nbr.assign((size_t)max_particles * MAXNBR, 0u);
  W->nbrN.assign(max_particles, 0u);
  W->bond.assign((size_t)max_particles * BOND_CAP, 0u);
  W->bondCand.assign((size_t)max_particles * BOND_CAP, 0u);
  return reinterpret_cast<or_world*>(W);
}

void or_destroy(or_world* w) { delete reinterpret_cast<World*>(w); }

void or_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

// Load a scene whose 8-word particle records were already built by the host
// exactly as loadScene builds them (src/engine.js:3888-3937).  Every buffer is
// cleared (src/engine.js:3865-3886), pred = pos for all n, the chain resets.
int or_load(or_world* w, uint32_t n, uint32_t nFluid, const int32_t* state8, int32_t domW,
            int32_t domH, uint32_t cellShift, uint32_t gridW, uint32_t gridH) {
  World& W = *reinterpret_cast<World*>(w);
  if (n > W.maxN || gridW * gridH > W.maxCells) return -1;
  W.n = n; W.nFluid = nFluid; W.domW = domW; W.domH = domH; W.cellShift = cellShift;
  W.gridW = gridW; W.gridH = gridH; W.cellTotal = gridW * gridH;
  W.substep = 0; W.bondDirty = true; W.parity = 0; W.weldHold = 0; W.fuseTag = 0;
  for (auto& s : W.state) std::fill(s.begin(), s.end(), Particle{});
  std::fill(W.derived.begin(), W.derived.end(), Derived{});
  std::fill(W.nbrN.begin(), W.nbrN.end(), 0u);
  std::fill(W.bond.begin(), W.bond.end(), 0u);
  std::fill(W.bondCand.begin(), W.bondCand.end(), 0u);
  for (auto& d : W.dbg) d.store(0);
  std::memcpy(W.state[0].data(), state8, (size_t)n * 32);
  std::memcpy(W.state[1].data(), state8, (size_t)n * 32);
  for (uint32_t i = 0; i < n; ++i) W.derived[i].pred = W.state[0][i].pos;
  for (int j = 0; j < 8; ++j) { W.chain[j] = CHAIN_INIT[j]; W.acc[j] = 0; }
  return 0;
}

// Advance `count` substeps; if acc_stream is given, store each substep's
// 8-word digest (H_set = words 0..3, H_slot = words 4..7).
void or_step(or_world* w, uint32_t count, uint32_t* acc_stream) {
  World& W = *reinterpret_cast<World*>(w);
  for (uint32_t s = 0; s < count; ++s) substep(W, acc_stream ? acc_stream + 8 * (size_t)s : nullptr);
}

void or_digest(or_world* w, uint32_t acc[8]) {
  World& W = *reinterpret_cast<World*>(w);
  digest(W, W.state[W.parity].data(), acc);
}
void or_chain(or_world* w, uint32_t out[8]) {
  World& W = *reinterpret_cast<World*>(w);
  for (int j = 0; j < 8; ++j) out[j] = W.chain[j];
}
void or_read_state(or_world* w, int32_t* out) {
  World& W = *reinterpret_cast<World*>(w);
  std::memcpy(out, W.state[W.parity].data(), (size_t)W.n * 32);
}
void or_read_derived(or_world* w, int32_t* out) {
  World& W = *reinterpret_cast<World*>(w);
  std::memcpy(out, W.derived.data(), (size_t)W.n * 64);
}
void or_read_bond(or_world* w, uint32_t* out) {
  World& W = *reinterpret_cast<World*>(w);
  std::memcpy(out, W.bond.data(), (size_t)W.n * BOND_CAP * 4);
}
// Exact restore of a snapshot: state into both parities, derived scratch back.
void or_write_state_derived(or_world* w, const int32_t* st, const int32_t* dv) {
  World& W = *reinterpret_cast<World*>(w);
  std::memcpy(W.state[0].data(), st, (size_t)W.n * 32);
  std::memcpy(W.state[1].data(), st, (size_t)W.n * 32);
  if (dv) std::memcpy(W.derived.data(), dv, (size_t)W.n * 64);
}
void or_debug_counters(or_world* w, uint32_t out[16]) {
  World& W = *reinterpret_cast<World*>(w);
  for (int j = 0; j < 16; ++j) out[j] = W.dbg[j].load();
}
uint32_t or_substep_index(or_world* w) { return reinterpret_cast<World*>(w)->substep; }

// Test probe: the first iteration of the next substep up to and including
// `friction`, then out = {sum dp.x, sum dp.y over fluid slots, neighbour
// entries, boundary entries, full lists}.  The world is left mid-substep
// (only for tests that discard it).
void or_friction_probe(or_world* w, int64_t out[5]) {
  World& W = *reinterpret_cast<World*>(w);
  g_dbg = W.dbg;
  const Particle* sin = W.state[W.parity].data();
  predict(W, sin);
  build_grid(W);
  build_nbr(W, sin);
  solve_a(W, sin);
  for (uint32_t c = 0; c < 4; ++c) {
    solve_b(W, sin, c);
    apply_dp(W, sin, c);
  }
  friction(W, sin);
  int64_t sx = 0, sy = 0, pairs = 0, bnd = 0, full = 0;
  for (uint32_t i = 0; i < W.nFluid; ++i) {
    sx += W.derived[i].dp.x;
    sy += W.derived[i].dp.y;
    const uint32_t cnt = nbr_count(W, i);
    pairs += cnt;
    full += cnt >= MAXNBR ? 1 : 0;
    for (uint32_t k = 0; k < cnt; ++k) bnd += W.nbr[(size_t)i * MAXNBR + k] >= W.nFluid ? 1 : 0;
  }
  out[0] = sx; out[1] = sy; out[2] = pairs; out[3] = bnd; out[4] = full;
}

// Test probe: the next substep up to its first friction pass (predict, grid,
// buildNbr, iteration 0's solveA and four colour passes, then friction with
// the bond gather, src/sim.wgsl:1673-1934), and each fluid slot's friction
// dp (2 words, ticks << ACC, before applyFric) and pred (2 words).  The world
// is left mid-substep (only for tests that discard it).
void or_friction_dp(or_world* w, int32_t* dp2, int32_t* pred2) {
  World& W = *reinterpret_cast<World*>(w);
  g_dbg = W.dbg;
  const Particle* sin = W.state[W.parity].data();
  predict(W, sin);
  build_grid(W);
  build_nbr(W, sin);
  solve_a(W, sin);
  for (uint32_t c = 0; c < 4; ++c) {
    solve_b(W, sin, c);
    apply_dp(W, sin, c);
  }
  friction(W, sin);
  for (uint32_t i = 0; i < W.nFluid; ++i) {
    dp2[2 * i] = W.derived[i].dp.x;
    dp2[2 * i + 1] = W.derived[i].dp.y;
    pred2[2 * i] = W.derived[i].pred.x;
    pred2[2 * i + 1] = W.derived[i].pred.y;
  }
}

// Test probe: the next substep through its solver iterations, then pred of
// every fluid slot (2 words) just before contain + containOverflow
// (src/sim.wgsl:2056-2235) and just after them.  The world is left
// mid-substep (only for tests that discard it).
void or_contain_probe(or_world* w, int32_t* before2, int32_t* after2) {
  World& W = *reinterpret_cast<World*>(w);
  g_dbg = W.dbg;
  const Particle* sin = W.state[W.parity].data();
  predict(W, sin);
  build_grid(W);
  build_nbr(W, sin);
  for (int it = 0; it < W.iters; ++it) {
    solve_a(W, sin);
    for (uint32_t c = 0; c < 4; ++c) {
      solve_b(W, sin, c);
      apply_dp(W, sin, c);
    }
    friction(W, sin);
    apply_fric(W, sin);
  }
  for (uint32_t i = 0; i < W.nFluid; ++i) {
    before2[2 * i] = W.derived[i].pred.x;
    before2[2 * i + 1] = W.derived[i].pred.y;
  }
  contain(W, sin);
  contain_overflow(W, sin);
  for (uint32_t i = 0; i < W.nFluid; ++i) {
    after2[2 * i] = W.derived[i].pred.x;
    after2[2 * i + 1] = W.derived[i].pred.y;
  }
}

// Test probe: the next substep's predict, grid and buildNbr only, then
// pred (2 words per slot), the neighbour lists (MAXNBR words per slot, in
// list order), nbrN (count | seal/ext-seal/congeal bits) and the truncation
// count buildNbr added to dbg[6].  The world is left mid-substep (only for
// tests that discard it).
void or_nbr_probe(or_world* w, int32_t* pred2, uint32_t* nbr48, uint32_t* nbrN, uint32_t* dropped) {
  World& W = *reinterpret_cast<World*>(w);
  g_dbg = W.dbg;
  const Particle* sin = W.state[W.parity].data();
  const uint32_t before = W.dbg[6].load();
  predict(W, sin);
  build_grid(W);
  build_nbr(W, sin);
  *dropped = W.dbg[6].load() - before;
  for (uint32_t i = 0; i < W.n; ++i) {
    pred2[2 * i] = W.derived[i].pred.x;
    pred2[2 * i + 1] = W.derived[i].pred.y;
    nbrN[i] = W.nbrN[i];
  }
  std::memcpy(nbr48, W.nbr.data(), (size_t)W.n * MAXNBR * 4);
}

// applyImpulse on the exact current state (src/engine.js:6077-6098): fluid
// slots in ascending order, JS-double circle test, up to `limit` hits, velocity
// VMAX-clamped.  Returns the hit count.
uint32_t or_apply_impulse(or_world* w, int32_t cx, int32_t cy, int32_t r, int32_t ix, int32_t iy,
                          uint32_t limit) {
  World& W = *reinterpret_cast<World*>(w);
  Particle* st = W.state[W.parity].data();
  const double r2 = (double)r * (double)r;
  uint32_t hits = 0;
  for (uint32_t i = 0; i < W.nFluid && hits < limit; ++i) {
    double dx = (double)st[i].pos.x - (double)cx;
    double dy = (double)st[i].pos.y - (double)cy;
    double q = dx * dx;
    q = q + dy * dy;
    if (q > r2) continue;
    int64_t vx = (int64_t)st[i].vel.x + ix, vy = (int64_t)st[i].vel.y + iy;
    vx = std::max<int64_t>(-65536, std::min<int64_t>(65536, vx));
    vy = std::max<int64_t>(-65536, std::min<int64_t>(65536, vy));
    st[i].vel.x = (int32_t)vx;
    st[i].vel.y = (int32_t)vy;
    hits++;
  }
  return hits;
}

// spawnFluid (src/engine.js:5901-6020): append k fluid records built by the
// host, shift the boundary block up by k slots (boundary records are given
// again by the host, rebuilt from the scene), pred = pos for the moved range,
// zero the new bond rows, mark formation dirty.  Returns the new n, or the old
// n when capacity would be exceeded.
uint32_t or_spawn(or_world* w, uint32_t k, const int32_t* fluid8, const int32_t* boundary8) {
  World& W = *reinterpret_cast<World*>(w);
  if (k == 0) return W.n;
  uint32_t nB = W.n - W.nFluid;
  uint32_t newFluid = W.nFluid + k, newN = newFluid + nB;
  if (newN > W.maxN) return W.n;
  Particle* st = W.state[W.parity].data();
  std::memcpy(&st[W.nFluid], fluid8, (size_t)k * 32);
  std::memcpy(&st[newFluid], boundary8, (size_t)nB * 32);
  for (uint32_t i = W.nFluid; i < newN; ++i) {
    Derived d{};
    d.pred = st[i].pos;
    W.derived[i] = d;
  }
  for (uint32_t i = W.nFluid; i < newFluid; ++i)
    for (uint32_t b = 0; b < BOND_CAP; ++b) {
      W.bond[(size_t)i * BOND_CAP + b] = 0u;
      W.bondCand[(size_t)i * BOND_CAP + b] = 0u;
    }
  W.n = newN;
  W.nFluid = newFluid;
  W.bondDirty = true;
  return W.n;
}

}  // extern "C"

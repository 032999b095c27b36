// MI355X-native deterministic PBF substep (2D) — kernels and C ABI.
//
// The reference (src/sim.wgsl, driven by src/engine.js:4003-4155) runs ~60
// dispatches per substep over particles in SLOT order and rebuilds a
// canonical neighbour order each substep.  This build keeps the particles
// permanently in CELL-SORTED order instead: every substep gathers the
// persistent state into the new (cell, slot-id) order, and all per-particle
// and pair passes run on that layout, so a particle's neighbours are its
// neighbours in memory.  The slot id travels with the particle and is what
// H_slot hashes, so results are independent of the memory order.
//
// Why this is bit-exact: every pair accumulator of the 2D solver is a plain
// wrapping i32 sum of per-term-clamped values (order-free), and the only
// order-sensitive step — which candidates survive the MAXNBR truncation in
// buildNbr (src/sim.wgsl:1190-1254) — walks candidates in exactly the
// reference's canonical order: stencil rows dy = -1..1, cells dx = -1..1,
// ascending slot id within a cell.  In sorted space that is three contiguous
// ranges, one per stencil row.
//
// Per substep (k density iterations), kernels in order:
//   predict_count  pred, cell, arrival rank           (src/sim.wgsl:958-1016)
//   scan x3        exclusive prefix of cell counts      (:1025-1106)
//   scatter        bucket = arrival order               (:1117-1124)
//   canon_gather   ascending-id rank within the cell, gather state (:1155-1168)
//   nbr_solveA     buildNbr fused with iteration 0's solveA (:1190-1318)
//   colour x4      solveB (k_colour) + applyDp (k_apply) of one colour, a
//                  workgroup / wave per cell chunk (:1337-1474); the last
//                  iteration also finalizes (:2245-2265)
//   solveA         iterations 1..k-1 (:1270-1318), then colour x4 again
//   xnc            xsph + normals + conduct (:2300-2582)
//   st_thermal     surfaceTension + thermal + formation flags + state hash
//                  (:2409-2707, :3039-3097, src/hash/statehash.wgsl:83-159)
//   hash_finish    chain fold (statehash.wgsl:164-170), H ring append
//
// Material passes (src/sim.wgsl:1673-2235, 2708-3097) run in worlds whose
// scenes hold fluid rows with friction or bonds, or that were created with
// materials = 1: friction + applyFric per iteration after the colour
// passes, contain before a separate finalize, and containStress, bond
// formation and a bond-folded hash after st_thermal.  Water-only worlds skip
// them (they are bit-inert for frictionless, bondless rows, :1698-1702,
// 2062, 2747, 2830); there a transition into a row with friction or bonds
// sets the scene's fault word and sn_sync reports SN_E_UNSUPPORTED.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sunaengine.h"
#include "fixed.hpp"
#include "scan.hpp"

using namespace sn;

namespace {

// ---- constants (src/sim.wgsl:55-171) --------------------------------------
constexpr int32_t H = 163840;
constexpr int32_t RHO0 = 1048576;
constexpr int32_t TWO_RHO0 = 2097152;
constexpr int32_t HS = 10240;
constexpr int32_t H2S = 104857600;
constexpr unsigned ACC = 8;
constexpr uint32_t MAXNBR = 48;
constexpr uint32_t TN = 3202;
constexpr int32_t LAM_MAX = 4194304;
constexpr int32_t S_TERM = 16777216;
constexpr int32_t DP_TERM = 2097152;
constexpr int32_t XS_TERM = 2097152;
constexpr int32_t ST_TERM = 4194304;
constexpr int32_t RINV_MAX = 262144;
constexpr int32_t WALL = 32768;
constexpr int32_t C_MAX = 58000;
constexpr int32_t RHO_FLOOR_RINV = 32768;
constexpr int32_t RHO_FLOOR_K = 65536;
constexpr int32_t AMBIENT = 1310720;
constexpr int32_t TEMP_MIN = -33554432, TEMP_MAX = 536870912;
constexpr int32_t DT_CLAMP = 16777216;
constexpr int32_t HT_TERM = 16777216;
constexpr int32_t CORR_TERM = 16777216;
constexpr int32_t CORR_NEVER = 2147483647;
constexpr uint32_t FLAG_BURNING = 1, FLAG_DEAD = 2, FLAG_ANCHOR = 4, FLAG_RIGID = 16,
                   FLAG_CONGEAL = 32, FLAG_BURST = 64, FLAG_SEAL = 128;
constexpr uint32_t WELD_TAG_MASK = 0x000fff00u;
constexpr uint32_t WELD_SEQ_MAX = 4095;   // 12-bit spawn-event ids, 0 = no tag (src/engine.js:268-275)
constexpr uint32_t NBR_SEAL_BIT = 0x80000000u, NBR_EXT_SEAL_BIT = 0x40000000u,
                   NBR_CONGEAL_BIT = 0x20000000u;
constexpr uint32_t NBR_BND = 0x80000000u;  // neighbour entry: j is a boundary slot
constexpr uint32_t MAT_WORDS = 48, MAT_MAX = 64;
constexpr uint32_t MAT_PHOTO_ROW = 15;
constexpr uint32_t BOND_PERIOD = 4;
constexpr uint32_t INACTIVE = 0xffffffffu;
constexpr uint32_t MIXED = 0xfffffffeu;     // a wave/block spanning scenes
constexpr int HPART_WORDS = 9;             // 8 hash words + truncation count
// per-particle info bits, rebuilt every substep by canon_gather
constexpr uint32_t INF_DEAD = 1, INF_FLUID = 2, INF_SEAL = 4, INF_CONGEAL = 8;
// scene fault bits
constexpr uint32_t FAULT_MATERIAL = 1;

enum MatW {
  M_EPS = 0, M_DPMAX, M_XSPH, M_PHANTOM, M_GCOH, M_GCUR, M_ADH, M_GRAV,
  M_COND, M_COOL, M_HEATCAP, M_MELTPT, M_MELTTO, M_FREEZEPT, M_FREEZETO,
  M_BOILPT, M_BOILTO, M_CONDPT, M_CONDTO, M_IGNITE, M_FUEL0, M_BURNRATE,
  M_BURNHEAT, M_BURNTO, M_SPLITTO, M_SPLITPCT, M_SPAWNT, M_TINT, M_CORRODE,
  M_CORRPT, M_CORRTO, M_FRIC, M_BONDK, M_BONDBREAK, M_BONDREFORM, M_BONDFORMV,
  M_BONDFORMP, M_BONDMU, M_BONDGROUP, M_RAMPLO, M_RAMPHI, M_RAMPINV,
  M_RAMPSIGN, M_SOLIDMODE, M_TEMPTARGET, M_TEMPDRIVE, M_BURSTP, M_SEALMODE
};

#define HIPCHECK(x)                                                              \
  do {                                                                           \
    hipError_t err_ = (x);                                                       \
    if (err_ != hipSuccess) {                                                    \
      throw DevError(std::string(#x) + ": " + hipGetErrorString(err_));          \
    }                                                                            \
  } while (0)

struct DevError {
  std::string msg;
  explicit DevError(std::string m) : msg(std::move(m)) {}
};

// ---- per-scene constants and dynamic state ---------------------------------
struct SceneP {
  uint32_t n, nFluid, gridW, gridH;
  uint32_t cellShift, cellBase, cellTotal, partBase;
  int32_t domW, domH, vmax, posShift;   // posShift = partBase - (particles of earlier scenes)
  uint32_t weldHold, cap, cellCap;
  uint32_t bounded;   // every position of the scene stays within +-2^28 (see pdelta) and the
                      // world's G table is below G_BOUND (see solveA_term)
};
struct SceneD {
  uint32_t substep;    // substeps encoded since load (`_substep`, src/engine.js:3811)
  uint32_t bondDirty;  // formation runs next substep regardless of the period
  uint32_t fault;
  uint32_t streamPos;  // entries written to this scene's H ring since reset
  uint32_t acc[8];
  uint32_t chain[8];
  uint32_t dbg[16];
  uint32_t evCursor;   // next unapplied stir event of this scene
  uint32_t evCount;    // events of this scene in the schedule
  uint32_t evBase;     // first event of this scene in the schedule
  uint32_t fired;      // substep of the event group being applied (bookkeeping)
  uint32_t flagged;    // nonzero once this substep's canonicalize met a SEAL or CONGEAL
                       // particle; cleared at the end of every substep
  uint32_t pushHits;   // particles moved by stir events and applyImpulse since load (wrapping)
};

struct Event {
  uint32_t scene, substep;
  int32_t cx, cy, r, ix, iy;
  uint32_t limit;
};

// Device pointers of one world, passed by value to every kernel.
struct Dev {
  // persistent state, current cell-sorted order
  int4* ps;        // pos.x, pos.y, vel.x, vel.y
  int4* pt;        // matId, flags, temp, fuel
  uint32_t* id;    // slot index within the scene
  int2* st;        // stAcc (Derived.stAcc, consumed by the next predict)
  // this substep's input snapshot, new order
  int4* psY;
  int4* ptY;
  uint32_t* idY;
  int2* stY;
  const uint32_t* sidOf;  // sorted position -> scene, INACTIVE if unused
  // temporaries
  int2* pred0;     // pred in the old order (predict)
  uint32_t* cell0; // global cell in the old order
  uint32_t* rank0; // arrival rank in the cell
  uint32_t* bucket;
  uint32_t* bucketId;
  uint32_t* cellCount;
  uint32_t* cellStart;
  uint32_t* blockSums;
  int2* pred;
  uint32_t* cellY;
  uint32_t* info;
  int32_t* rho;
  int32_t* lam;
  int4* pl4;       // {pred.x, pred.y, lam, 0} per slot: the colour pass's one gather per neighbour
  int4* vr4;       // {vtmp.x, vtmp.y, rho, rinv} per fluid slot after finalize: xnc's one gather for the three
  int4* nr4;       // {nrm.x, nrm.y, rho, material} per fluid slot after xnc: surface tension's one gather
  int2* dp;
  int2* vtmp;
  int32_t* rinv;
  int2* nrm;
  int32_t* heat;
  int32_t* corr;
  uint32_t* nbr;
  uint32_t* nbrN;
  uint32_t* drop;         // candidates dropped by truncation, per particle
  uint32_t* crank;        // partial canonical ranks of heavy-cell slots (zero between uses)
  uint32_t* hperm;        // packed one-cell tiles: their slots in spatial (Morton bin) order
  uint32_t* cellBin;      // per substep: cell -> its bin table slot (BIN_NONE: the cell is not binned)
  uint4* binTab;          // [slot][64] per Morton bin {first, end (offsets into hperm from the cell's
                          // first slot), live (not DEAD) slots, 0}
  int4* binBox;           // [slot][64] box of the live slots' predicted positions (x0, y0, x1, y1)
  uint32_t binCap;        // slots of binTab / binBox
  int4* blkBox;           // per substep: box of the live positions of each aligned 64-slot block
  const uint32_t* blockMap;  // per-particle kernels: logical block -> 256-slot block holding live slots
  uint32_t* hpart;        // per-block partial sums [block][HPART_WORDS]
  const uint32_t* cellScene;  // global cell -> scene, INACTIVE if its scene is inactive
  const uint2* tileAll;   // every neighbour tile of the active scenes {first cell, ncell | scene << 8}
  uint2* dlist[7];        // this substep's occupied cells per colour [0..3], occupied tiles [4],
                          // chunks of heavy cells for the canonical rank [5], tiles with
                          // heavy candidate rows [6]
  uint32_t* listCount;    // [0..6] entries of dlist, [7] next item of dlist[6], [8] bin table slots
                          // taken (cleared by predict)
  uint32_t nTiles;
  uint32_t dbgOn;         // 0: the reference's debug:false build, counters stay zero
  const int4* lut4;       // {W, G, C, S} per LUT bin
  const int32_t* lutA;    // adhesion table
  const int32_t* mats;
  const SceneP* sp;
  SceneD* sd;
  uint32_t* hring;        // [S][ringCap][8]
  uint32_t* hacc;         // scratch accumulators for digest-only [S][8]
  const Event* events;
  uint32_t* hitFlag;      // slot-ordered hit flags for stir events
  uint32_t* hitScan;
  // worlds that run every material pass (gen = 1): bond rows in slot order,
  // row (partBase + slot) of BOND_CAP words, and the sorted position of each
  // slot this substep, so a bond word's partner slot finds its state
  uint32_t* bond;
  uint32_t* bondCand;     // bondForm1's candidate rows, same layout
  uint32_t* posOf;        // partBase + slot -> sorted position (rebuilt by canon_place)
  uint32_t gen;
  uint64_t sealRows;      // bit m: material row m is seal class
  uint32_t N, C, S, ringCap;
  uint32_t nEvents;
  uint32_t NS;  // neighbour-list stride: nbr[k * NS + p]
};

__device__ __forceinline__ const int32_t* MAT(const Dev& D, uint32_t m) { return D.mats + m * MAT_WORDS; }

// ---- shared helpers (src/sim.wgsl:477-594) -------------------------------
struct CC { int32_t x, y; };
__device__ __forceinline__ CC cell_coord(const SceneP& P, int2 p) {
  int32_t cx = p.x >> P.cellShift;   // originX = originY = 0
  int32_t cy = p.y >> P.cellShift;
  return CC{clampi(cx, 0, (int32_t)P.gridW - 1), clampi(cy, 0, (int32_t)P.gridH - 1)};
}
__device__ __forceinline__ int32_t lut_index(int2 d, Dbg& g) {
  int32_t ax = wadd(fp_abs(d.x, g), 8) >> 4;
  int32_t ay = wadd(fp_abs(d.y, g), 8) >> 4;
  if (ax > HS || ay > HS) return -1;
  int32_t r2 = ax * ax + ay * ay;
  if (r2 >= H2S) return -1;
  return r2 >> 15;
}
__device__ __forceinline__ int2 pred_delta(int2 pi, int2 pj, Dbg& g) {
  return make_int2(fp_sub_sat(pi.x, pj.x, g), fp_sub_sat(pi.y, pj.y, g));
}
// Pair geometry.  In a bounded scene every position lies within +-2^28: the
// domain and every loaded or spawned position are checked on the host, fluid
// predictions are wall-clamped into the domain and boundary slots never
// move.  A difference of two positions then never overflows and its absolute
// value is exact, so the saturating forms above reduce to plain arithmetic
// and count no event; and with ax, ay <= HS the squared radius fits the
// 24-bit multiplier.  Scenes that are not bounded take the exact saturating
// forms.
constexpr int32_t POS_BOUND = 1 << 28;
template <bool BND>
__device__ __forceinline__ int2 pdelta(int2 pi, int2 pj, Dbg& g) {
  if constexpr (BND) return make_int2(pi.x - pj.x, pi.y - pj.y);
  else return pred_delta(pi, pj, g);
}
template <bool BND>
__device__ __forceinline__ int32_t lut_idx(int2 d, Dbg& g) {
  if constexpr (BND) {
    const uint32_t ax = ((uint32_t)abs(d.x) + 8u) >> 4, ay = ((uint32_t)abs(d.y) + 8u) >> 4;
    // clamped at HS + 1, an axis past HS alone gives r2 >= (HS + 1)^2 >= H2S
    const uint32_t axc = min(ax, (uint32_t)HS + 1u), ayc = min(ay, (uint32_t)HS + 1u);
    const uint32_t r2 = __umul24(axc, axc) + __umul24(ayc, ayc);
    return r2 >= (uint32_t)H2S ? -1 : (int32_t)(r2 >> 15);
  } else {
    return lut_index(d, g);
  }
}
// lut_idx for a bounded scene when |d| < 2^19 on both axes (a candidate of a
// chunk of one cell no wider than 2^18, inside the chunk's reach box, is
// within 2^18 + REACH): then ax, ay < 2^15, the
// squared radius fits 32 bits, and ax > HS already gives r2 >= (HS + 1)^2 >
// H2S, so the radius test alone decides.
__device__ __forceinline__ int32_t lut_idx_near(int2 d) {
  const uint32_t ax = ((uint32_t)abs(d.x) + 8u) >> 4, ay = ((uint32_t)abs(d.y) + 8u) >> 4;
  const uint32_t r2 = __umul24(ax, ax) + __umul24(ay, ay);
  return r2 >= (uint32_t)H2S ? -1 : (int32_t)(r2 >> 15);
}
static_assert((uint64_t)(HS + 1) * (HS + 1) >= (uint64_t)H2S, "ax > HS implies r2 >= H2S");
__device__ __forceinline__ int32_t grad_q16(int32_t gt, int32_t d, Dbg& g) { return fp_mul_s(gt, d, 26, g); }
__device__ __forceinline__ int2 wall_clamp(const SceneP& P, int2 p) {
  return make_int2(clampi(p.x, WALL, P.domW - WALL), clampi(p.y, WALL, P.domH - WALL));
}
__device__ __forceinline__ int32_t mat_mix(int32_t a, int32_t b) { return wadd(a, b) >> 1; }
__device__ __forceinline__ int32_t sym_k(int32_t rhoi, int32_t rhoj, Dbg& g) {
  int32_t den = maxi(RHO_FLOOR_K, fp_add_sat(rhoi, rhoj, g));
  return clampi(fp_divshift(TWO_RHO0, den, 16, g), 0, RINV_MAX);
}
__device__ __forceinline__ uint32_t id_hash(uint32_t i) {
  uint32_t x = i + 0x9e3779b9u;
  x = (x ^ (x >> 16)) * 0x21f0aaadu;
  x = (x ^ (x >> 15)) * 0x735a2d97u;
  return x ^ (x >> 15);
}
__device__ __forceinline__ uint32_t pick_product(uint32_t primary, uint32_t splitTo, uint32_t pct, uint32_t i) {
  if (pct == 0u) return primary;
  if ((id_hash(i) & 255u) < pct) return splitTo;
  return primary;
}
__device__ __forceinline__ bool seal_class(const int32_t* m) {
  return m[M_BONDK] != 0 && m[M_BONDREFORM] == 0 && m[M_SOLIDMODE] != 0 && m[M_PHANTOM] == 0;
}
__device__ __forceinline__ bool fuse_row(const int32_t* m, uint32_t id) {
  return m[M_BONDK] != 0 && m[M_BONDREFORM] == 0 && m[M_BONDGROUP] >= 0 && id != MAT_PHOTO_ROW;
}
// Rows this build advances: every term they take part in is one of the
// kernels above (frictionless, bondless, not a seal endpoint).
__device__ __forceinline__ bool supported_row(const int32_t* m) {
  return m[M_FRIC] == 0 && m[M_BONDK] == 0;
}

// Wrapping u32 sum over the 64 lanes (every lane must be active).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
  return v;
}

// Fold the per-thread debug counters into the scene's dbg words: the wave
// sums each nonzero counter first and issues one atomic for it when the wave
// is uniform in scene (the counters are order-free u32 sums).  Must be called
// with every lane of the wave active.
__device__ __forceinline__ void flush_dbg(const Dev& D, uint32_t s, const Dbg& g, uint32_t trunc) {
  if (!D.dbgOn) return;
  uint32_t v[5] = {g.mul, g.div, g.div0, g.add, trunc};
  const uint32_t slot[5] = {0, 1, 2, 5, 6};
  const uint32_t s0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)s);
  const bool uniform = __all(s == s0);
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    uint32_t x = v[k];
    if (!__any(x != 0)) continue;
    if (uniform) {
      x = wave_sum(x);
      if ((threadIdx.x & 63u) == 0u && s0 != INACTIVE) atomicAdd(&D.sd[s0].dbg[slot[k]], x);
    } else if (s != INACTIVE && x) {
      atomicAdd(&D.sd[s].dbg[slot[k]], x);
    }
  }
}

// Workgroup b is dispatched to XCD b % 8 (8 XCDs, each with its own L2).
// Map hardware blocks so that each XCD walks one contiguous run of logical
// blocks — a contiguous slab of the cell-sorted particle order — and its L2
// holds that slab's neighbourhood.  A bijection on [0, gridDim.x).
__device__ __forceinline__ uint32_t xcd_block() {
  uint32_t b = blockIdx.x, nb = gridDim.x;
  uint32_t full = nb & ~7u;
  if (b >= full) return b;
  return (b & 7u) * (full >> 3) + (b >> 3);
}

// The 256-slot block a per-particle kernel's workgroup serves.  Those
// kernels are launched over exactly the blocks that hold live slots of
// active scenes (blockMap, rebuilt with the layout), so the idle capacity of
// a slotted world costs no workgroups; XCD order is kept through xcd_block.
__device__ __forceinline__ uint32_t slot_block(const Dev& D) {
  return D.blockMap[xcd_block()];
}

// ===================================================================== predict
// predict + gridCount (src/sim.wgsl:958-1016) over the old order; the cell
// counter's atomic return value is the particle's arrival rank in its cell,
// which replaces the reference's second atomic pass in `scatter`.
__global__ __launch_bounds__(256) void k_predict_count(Dev D) {
  uint32_t p = slot_block(D) * 256u + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x < 9u) D.listCount[threadIdx.x] = 0u;  // rebuilt after the scan
  uint32_t s = p < D.N ? D.sidOf[p] : INACTIVE;
  Dbg g;
  uint32_t cell = INACTIVE;
  if (s != INACTIVE) {
    const SceneP& P = D.sp[s];
    int4 a = D.ps[p];
    int4 b = D.pt[p];
    uint32_t id = D.id[p];
    int2 pos = make_int2(a.x, a.y);
    int2 pr = pos;
    if (id < P.nFluid) {
      const int32_t* M = MAT(D, (uint32_t)b.x);
      uint32_t fl = (uint32_t)b.y;
      if (M[M_PHANTOM] == 0 && (fl & FLAG_ANCHOR) == 0u) {
        int2 st = D.st[p];
        int32_t vmaxN = fp_neg_sat(P.vmax, g);
        int32_t vx = fp_add_sat(fp_add_sat(a.z, st.x, g), 0, g);
        int32_t vy = fp_add_sat(fp_add_sat(fp_add_sat(a.w, st.y, g), M[M_GRAV], g), 0, g);
        vx = clampi(vx, vmaxN, P.vmax);
        vy = clampi(vy, vmaxN, P.vmax);
        pr = wall_clamp(P, make_int2(fp_add_sat(pos.x, vx, g), fp_add_sat(pos.y, vy, g)));
      }
    }
    CC c = cell_coord(P, pr);
    cell = P.cellBase + (uint32_t)c.y * P.gridW + (uint32_t)c.x;
    D.pred0[p] = pr;
    D.cell0[p] = cell;
  }
  // Arrival ranks (wg_cell_rank): the old order is sorted by last substep's
  // cell, but in a moving fluid half the particles change cell each
  // substep, so a wave holds about 38 runs of equal new cells while a
  // workgroup's 256 slots land in about 15 cells.  Any assignment of
  // distinct ranks within a cell is valid (the canonical gather re-ranks by
  // slot id).
  {
    const uint32_t r = wg_cell_rank(D.cellCount, cell, INACTIVE);
    if (cell != INACTIVE) D.rank0[p] = r;
  }
  flush_dbg(D, s, g, 0);
}

// ======================================================================= scan
// cell-count prefix: see scan.hpp

// ==================================================================== scatter
__global__ __launch_bounds__(256) void k_scatter(Dev D) {
  uint32_t p = slot_block(D) * 256u + threadIdx.x;
  if (p >= D.N) return;
  uint32_t s = D.sidOf[p];
  if (s == INACTIVE) return;
  int32_t shift = D.sp[s].posShift;
  uint32_t c = D.cell0[p];
  uint32_t q = (uint32_t)((int32_t)D.cellStart[c] + shift) + D.rank0[p];
  D.bucket[q] = p;
  D.bucketId[q] = D.id[p];
}

// =============================================================== canon_gather
// canonicalize (src/sim.wgsl:1155-1168): rank = #{ids in my cell below mine};
// the particle's new position is cellStart + rank.  The whole persistent
// record moves there, together with pred and the cell.
//
// A cell of up to CANON_HEAVY particles is ranked by each thread walking the
// cell.  Walls and collisions can pack thousands into one cell, where that
// walk is quadratic on a handful of waves; such cells are listed per substep
// (dlist[5], CANON_ITEM particles per item) and k_canon_heavy ranks them with
// the cell's ids streamed through LDS, spread over the whole chip.
constexpr uint32_t CANON_HEAVY = 256;
constexpr uint32_t CANON_ITEM = 256;
constexpr uint32_t CANON_TILE = 2048;
constexpr uint32_t CANON_SPLIT = 16;      // id-range parts per heavy item, at most
constexpr uint32_t CANON_PART_MIN = 512;  // ids per part, at least

__device__ __forceinline__ void canon_place(const Dev& D, uint32_t s, uint32_t me, uint32_t myid, uint32_t c,
                                            uint32_t q) {
  const SceneP& P = D.sp[s];
  int4 pt = D.pt[me];
  D.psY[q] = D.ps[me];
  D.ptY[q] = pt;
  D.idY[q] = myid;
  D.stY[q] = D.st[me];
  const int2 pr = D.pred0[me];
  D.pred[q] = pr;
  D.pl4[q] = make_int4(pr.x, pr.y, 0, 0);
  D.cellY[q] = c;
  uint32_t fl = (uint32_t)pt.y;
  bool fluid = myid < P.nFluid;
  const uint32_t inf = ((fl & FLAG_DEAD) ? INF_DEAD : 0u) | (fluid ? INF_FLUID : 0u) |
                       ((fluid && (fl & FLAG_SEAL)) ? INF_SEAL : 0u) |
                       ((fluid && (fl & FLAG_CONGEAL)) ? INF_CONGEAL : 0u);
  D.info[q] = inf;
  if (inf & (INF_SEAL | INF_CONGEAL)) atomicOr(&D.sd[s].flagged, 1u);
  if (D.gen) D.posOf[P.partBase + myid] = q;
}

__global__ __launch_bounds__(256) void k_canon_gather(Dev D) {
  uint32_t t = slot_block(D) * 256u + threadIdx.x;
  if (t >= D.N) return;
  uint32_t s = D.sidOf[t];
  if (s == INACTIVE) return;
  const SceneP& P = D.sp[s];
  uint32_t me = D.bucket[t];
  uint32_t myid = D.bucketId[t];
  uint32_t c = D.cell0[me];
  uint32_t lo = (uint32_t)((int32_t)D.cellStart[c] + P.posShift);
  uint32_t hi = (uint32_t)((int32_t)D.cellStart[c + 1] + P.posShift);
  if (hi - lo > CANON_HEAVY) return;   // k_canon_heavy_rank / _place
  uint32_t rank = 0;
  for (uint32_t u = lo; u < hi; ++u) rank += (D.bucketId[u] < myid) ? 1u : 0u;
  canon_place(D, s, me, myid, c, lo + rank);
}

// Heavy cells, in two passes.  Rank: work item (i, part) counts, for the
// CANON_ITEM slots of heavy item i, the ids below theirs in one of up to
// CANON_SPLIT parts of the cell, streamed through LDS in tiles, and adds the
// count to crank; a cell of k slots is thus spread over about
// k / CANON_ITEM * min(16, k / 512) workgroups.  Place: moves each slot to
// lo + its rank and clears crank for the next substep.  Ids are distinct
// within a scene and below 2^32 - 1, so the all-ones padding never counts.
struct HeavySpan {
  uint32_t c, s, lo, hi, t;
};
__device__ __forceinline__ HeavySpan heavy_item(const Dev& D, uint32_t i) {
  const uint2 item = D.dlist[5][i];
  HeavySpan h;
  h.c = (uint32_t)__builtin_amdgcn_readfirstlane((int)item.x);
  const uint32_t chunk = (uint32_t)__builtin_amdgcn_readfirstlane((int)item.y);
  h.s = D.cellScene[h.c];
  const int32_t sh = D.sp[h.s].posShift;
  h.lo = (uint32_t)((int32_t)D.cellStart[h.c] + sh);
  h.hi = (uint32_t)((int32_t)D.cellStart[h.c + 1] + sh);
  h.t = h.lo + chunk * CANON_ITEM + threadIdx.x;
  return h;
}

// Heavy cells of up to SORT_MAX slots: one workgroup per cell sorts the
// cell's slot ids in LDS (bitonic network over the next power of two, pads
// sort last) and gives each slot the position of its id in the sorted run,
// i.e. #{ids below it}: O(k log k) instead of the pairwise count's O(k^2).
// Ids are distinct, so the position is exact.  Larger cells take the
// pairwise count below.
constexpr uint32_t SORT_MAX = 16384;
constexpr uint32_t SORT_THREADS = 1024;

__global__ __launch_bounds__(SORT_THREADS) void k_canon_heavy_sort(Dev D) {
  __shared__ uint32_t key[SORT_MAX];
  const uint32_t n = (uint32_t)__builtin_amdgcn_readfirstlane((int)D.listCount[5]);
  for (uint32_t w = blockIdx.x; w < n; w += gridDim.x) {
    const uint2 item = D.dlist[5][w];
    if (__builtin_amdgcn_readfirstlane((int)item.y) != 0) continue;   // the cell's first item only
    const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)item.x);
    const int32_t sh = D.sp[D.cellScene[c]].posShift;
    const uint32_t lo = (uint32_t)((int32_t)D.cellStart[c] + sh);
    const uint32_t k = (uint32_t)((int32_t)D.cellStart[c + 1] + sh) - lo;
    if (k > SORT_MAX) continue;
    uint32_t P = 1;
    while (P < k) P <<= 1;
    __syncthreads();   // the previous cell's keys are no longer read
    for (uint32_t i = threadIdx.x; i < P; i += SORT_THREADS) key[i] = i < k ? D.bucketId[lo + i] : 0xffffffffu;
    __syncthreads();
    for (uint32_t size = 2; size <= P; size <<= 1) {
      for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
        for (uint32_t i = threadIdx.x; i < (P >> 1); i += SORT_THREADS) {
          const uint32_t a = 2u * i - (i & (stride - 1u)), b = a + stride;
          const uint32_t x = key[a], y = key[b];
          if ((x > y) == ((a & size) == 0u)) {
            key[a] = y;
            key[b] = x;
          }
        }
        __syncthreads();
      }
    }
    for (uint32_t i = threadIdx.x; i < k; i += SORT_THREADS) {
      const uint32_t id = D.bucketId[lo + i];
      uint32_t l = 0, r = k;
      while (l < r) {
        const uint32_t m = (l + r) >> 1;
        if (key[m] < id) l = m + 1u;
        else r = m;
      }
      D.crank[lo + i] = l;
    }
  }
}

__global__ __launch_bounds__(256) void k_canon_heavy_rank(Dev D) {
  __shared__ uint4 ids[CANON_TILE / 4];
  const uint32_t n = (uint32_t)__builtin_amdgcn_readfirstlane((int)D.listCount[5]) * CANON_SPLIT;
  for (uint32_t w = blockIdx.x; w < n; w += gridDim.x) {
    const HeavySpan h = heavy_item(D, w / CANON_SPLIT);
    const uint32_t part = w % CANON_SPLIT, k = h.hi - h.lo;
    if (k <= SORT_MAX) continue;   // ranked by k_canon_heavy_sort (uniform over the block)
    const uint32_t nparts = min(CANON_SPLIT, (k + CANON_PART_MIN - 1u) / CANON_PART_MIN);
    if (part >= nparts) continue;   // uniform over the block
    const uint32_t span = (k + nparts - 1u) / nparts;
    const uint32_t ilo = h.lo + part * span, ihi = min(h.hi, ilo + span);
    const bool valid = h.t < h.hi;
    const uint32_t myid = valid ? D.bucketId[h.t] : 0u;
    uint32_t rank = 0;
    for (uint32_t base = ilo; base < ihi; base += CANON_TILE) {
      __syncthreads();   // the previous tile is no longer read
      uint32_t* flat = reinterpret_cast<uint32_t*>(ids);
      for (uint32_t i = threadIdx.x; i < CANON_TILE; i += 256u)
        flat[i] = base + i < ihi ? D.bucketId[base + i] : 0xffffffffu;
      __syncthreads();
      const uint32_t nq = (min(CANON_TILE, ihi - base) + 3u) / 4u;
#pragma unroll 4
      for (uint32_t i = 0; i < nq; ++i) {
        const uint4 v = ids[i];
        rank += (v.x < myid ? 1u : 0u) + (v.y < myid ? 1u : 0u) + (v.z < myid ? 1u : 0u) + (v.w < myid ? 1u : 0u);
      }
    }
    if (valid && rank) atomicAdd(&D.crank[h.t], rank);
  }
}

__global__ __launch_bounds__(256) void k_canon_heavy_place(Dev D) {
  const uint32_t n = (uint32_t)__builtin_amdgcn_readfirstlane((int)D.listCount[5]);
  for (uint32_t w = blockIdx.x; w < n; w += gridDim.x) {
    const HeavySpan h = heavy_item(D, w);
    if (h.t >= h.hi) continue;
    const uint32_t r = D.crank[h.t];
    D.crank[h.t] = 0u;
    canon_place(D, h.s, D.bucket[h.t], D.bucketId[h.t], h.c, h.lo + r);
  }
}

// ===================================================== buildNbr + solveA(it 0)
// Accumulate one solveA term (src/sim.wgsl:1281-1301).
struct SolveAAcc {
  int32_t rho, S, sgx, sgy;
};
// Bounded pairs (BND) take the binary64 forms of fixed.hpp: |d| <= REACH <
// 2^17.33 per axis and |G| < 2^27 (G_BOUND; pair_bounds checks the world's
// tables and materials, and a world that fails runs every scene unbounded), so
// |G d| < 2^44.33 and each gradient is exact with |g| < 2^18.33 (no
// saturation); g^2 / 2^11 < 2^25.7 is exact and below the i32 rails, so the
// S terms only take their [0, S_TERM] clamp (g^2 >= 0).
constexpr int32_t G_BOUND = 1 << 27;
template <bool BND>
__device__ __forceinline__ void solveA_term(SolveAAcc& a, int2 d, int4 L, Dbg& g) {
  a.rho = wadd(a.rho, L.x);
  if constexpr (BND) {
    const double gs = (double)L.y * 0x1p-26;
    const double gx = q_rne(gs * (double)d.x), gy = q_rne(gs * (double)d.y);
    a.sgx = wadd(a.sgx, (int32_t)gx);
    a.sgy = wadd(a.sgy, (int32_t)gy);
    a.S = wadd(wadd(a.S, (int32_t)__builtin_fmin(q_rne(gx * gx * 0x1p-11), (double)S_TERM)),
               (int32_t)__builtin_fmin(q_rne(gy * gy * 0x1p-11), (double)S_TERM));
  } else {
    int32_t gx = grad_q16(L.y, d.x, g), gy = grad_q16(L.y, d.y, g);
    a.sgx = wadd(a.sgx, gx);
    a.sgy = wadd(a.sgy, gy);
    a.S = wadd(wadd(a.S, clampi(fp_mul_s(gx, gx, 11, g), 0, S_TERM)), clampi(fp_mul_s(gy, gy, 11, g), 0, S_TERM));
  }
}
__device__ __forceinline__ void solveA_finish(const Dev& D, uint32_t p, int2 pi, SolveAAcc& a, int32_t eps,
                                              Dbg& g) {
  a.S = wadd(wadd(a.S, clampi(fp_mul_s(a.sgx, a.sgx, 11, g), 0, S_TERM)),
             clampi(fp_mul_s(a.sgy, a.sgy, 11, g), 0, S_TERM));
  int32_t Cq = clampi(fp_shr_rne(wsub(a.rho, RHO0), 4), 0, C_MAX);
  D.rho[p] = a.rho;
  const int32_t lam = clampi(fp_neg_sat(fp_divshift(Cq, wadd(a.S, eps), 25, g), g), -LAM_MAX, LAM_MAX);
  D.lam[p] = lam;
  D.pl4[p] = make_int4(pi.x, pi.y, lam, 0);   // pi: the particle's current pred
}

// ============================================================ neighbour loops
// Pair loops walk a particle's list NBR_UNROLL entries at a time with every
// load of the group issued before any term is formed, so each lane keeps
// several gathers in flight.  Entries past the count read as the particle
// itself (a zero offset, no saturation) and are skipped; every pair sum is a
// wrapping u32 sum, so grouping does not change it.
constexpr uint32_t NBR_UNROLL = 4;
constexpr uint32_t XNC_UNROLL = 2;
static_assert(MAXNBR % NBR_UNROLL == 0 && MAXNBR % XNC_UNROLL == 0, "list slots k + u stay below MAXNBR");

// ============================================================ per-substep work
// After the scan: work lists of this substep.  [0..3]: every occupied cell
// of colour c, one item per CELL_CHUNK of its particles, resolved to {first
// sorted slot, count | scene << COL_SCENE_SHIFT}; [4]:
// every occupied neighbour tile, one item {tile, chunk} per TILE_CHUNK of its
// particles.  Chunking bounds every item's work, so a cell that walls or
// collisions have packed with hundreds of particles is spread over many
// waves instead of stalling one.  Items are appended with one atomic per
// workgroup and list; every item is processed independently of the others, so list
// order (which the atomics leave open) affects no result.
constexpr int NBR_PPL = 1;             // particles per lane in the neighbour sweep
constexpr uint32_t CELL_CHUNK = 32;
constexpr uint32_t COLOUR_THREADS = 256;   // k_colour: 8 lanes for each particle of a 32-particle item
constexpr uint32_t COL_COUNT = 0xffu, COL_SCENE_SHIFT = 8;
constexpr uint32_t TILE_CHUNK = 64;    // 64 lanes x NBR_PPL
// A tile whose candidate rows hold more than this many slots (a pile-up in
// or next to it; 576 at 4x rest density) goes to dlist[6], whose chunks every
// wave of the sweep takes first, one at a time from a shared counter, so that
// such work spreads over the whole chip instead of one XCD's list range.
// On the 1M dam break: 1024..2048 measure alike, 4096 is 2% slower, 8192 50%.
constexpr uint32_t NBR_HEAVY_CAND = 2048;
constexpr uint32_t BIN_NONE = 0xffffffffu;   // cellBin: the cell has no bin table this substep

// Called with every lane of every wave active.
__global__ __launch_bounds__(256) void k_build_lists(Dev D) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  {
    uint32_t cnt = 0, col = 0, lo = 0, s = 0;
    if (t < D.C) {
      D.cellBin[t] = BIN_NONE;   // k_heavy_order bins this substep's packed cells
      s = D.cellScene[t];
      if (s != INACTIVE) {
        cnt = D.cellStart[t + 1] - D.cellStart[t];
        const SceneP& P = D.sp[s];
        const uint32_t lc = t - P.cellBase;
        col = ((lc % P.gridW) & 1u) | (((lc / P.gridW) & 1u) << 1);
        lo = (uint32_t)((int32_t)D.cellStart[t] + P.posShift);
      }
    }
    const uint32_t nch = (cnt + CELL_CHUNK - 1u) / CELL_CHUNK;
#pragma unroll
    for (uint32_t c = 0; c < 4; ++c) {
      const uint32_t base = wg_reserve(col == c ? nch : 0u, &D.listCount[c]);
      if (col == c)
        for (uint32_t k = 0; k < nch; ++k)
          D.dlist[c][base + k] =
              make_uint2(lo + k * CELL_CHUNK, min(CELL_CHUNK, cnt - k * CELL_CHUNK) | (s << COL_SCENE_SHIFT));
    }
  }
  {
    uint32_t nch = 0;
    bool heavy = false;
    if (t < D.nTiles) {
      const uint2 it = D.tileAll[t];
      const uint32_t c0 = it.x, ncell = it.y & 0xffu;
      const uint32_t cnt = D.cellStart[c0 + ncell] - D.cellStart[c0];
      nch = (cnt + TILE_CHUNK - 1u) / TILE_CHUNK;
      if (nch) {
        // the tile's candidates: its three rows over x in cx0-1 .. cx0+ncell
        const SceneP& P = D.sp[it.y >> 8];
        const uint32_t lc0 = c0 - P.cellBase;
        const int32_t cy = (int32_t)(lc0 / P.gridW), cx0 = (int32_t)(lc0 % P.gridW);
        const int32_t xlo = maxi(cx0 - 1, 0), xhi = mini(cx0 + (int32_t)ncell, (int32_t)P.gridW - 1);
        uint32_t cand = 0;
        for (int32_t gy = maxi(cy - 1, 0); gy <= mini(cy + 1, (int32_t)P.gridH - 1); ++gy) {
          const uint32_t rowc = P.cellBase + (uint32_t)gy * P.gridW;
          cand += D.cellStart[rowc + (uint32_t)xhi + 1u] - D.cellStart[rowc + (uint32_t)xlo];
        }
        heavy = cand > NBR_HEAVY_CAND;
      }
    }
    wg_append_chunks(heavy ? 0u : nch, &D.listCount[4], D.dlist[4], t);
    wg_append_chunks(heavy ? nch : 0u, &D.listCount[6], D.dlist[6], t);
  }
  {
    uint32_t nch = 0;
    if (t < D.C && D.cellScene[t] != INACTIVE) {
      const uint32_t cnt = D.cellStart[t + 1] - D.cellStart[t];
      if (cnt > CANON_HEAVY) nch = (cnt + CANON_ITEM - 1u) / CANON_ITEM;
    }
    wg_append_chunks(nch, &D.listCount[5], D.dlist[5], t);
  }
}

// Persistent kernels walk a device list: XCD x (= blockIdx.x % 8) takes the
// x-th eighth of the list, so the items one XCD's L2 serves stay together.
struct ListRange {
  uint32_t lo, hi, first, stride;
};
__device__ __forceinline__ ListRange xcd_range(uint32_t n, uint32_t unitsPerBlock, uint32_t unit) {
  const uint32_t xcd = blockIdx.x & 7u, nb = gridDim.x >> 3, lb = blockIdx.x >> 3;
  ListRange r;
  r.lo = (uint32_t)(((uint64_t)n * xcd) >> 3);
  r.hi = (uint32_t)(((uint64_t)n * (xcd + 1u)) >> 3);
  r.first = r.lo + lb * unitsPerBlock + unit;
  r.stride = nb * unitsPerBlock;
  return r;
}

// One wave per occupied tile of up to 8 consecutive cells of one row; up to
// NBR_PPL particles per lane, so a tile of up to 64 * NBR_PPL particles is
// served by one sweep.  The tile's candidates — the union of its particles'
// 3x3 stencils: rows cy-1..cy+1, columns x0-1..x1+1, each row one contiguous
// sorted range — are loaded 64 at a time, one per lane, and broadcast to
// every lane with v_readlane, so the candidate loop issues no per-lane
// loads.  Walking rows in order and each row in sorted order visits every
// particle's own candidates in exactly its canonical order (dy outer, dx
// inner, ascending id), which is what decides truncation.  Iteration 0 of
// solveA is formed in the same sweep from (W, G) staged in LDS.
// One chunk of one tile: the sweep described above.  BND selects the bounded
// pair geometry; WIN is false for one-cell tiles, whose candidate union is
// exactly every particle's own 3x3 stencil, so no per-lane column window
// applies; PPL is the particles per lane (1 when the chunk fits one wave).
// WIN false is used only for one-cell tiles of cells no wider than 2^18.
// In a bounded scene a candidate is accepted only with |dx|, |dy| <= REACH
// (lut_idx rejects (|d| + 8) >> 4 > HS), so a candidate outside the chunk's
// bounding box grown by REACH is rejected by every lane and has no other
// effect; each block of 64 candidates is reduced to a ballot of the ones
// inside that box and the sweep visits only those, still in ascending order.
constexpr int32_t REACH = (HS + 1) * 16 - 9;   // largest |d| with (|d| + 8) >> 4 <= HS

__device__ __forceinline__ int32_t wave_min(int32_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = mini(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ int32_t wave_max(int32_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = maxi(v, __shfl_xor(v, off));
  return v;
}

// The part of a stencil row (row gy, cells xa..xb) that a reach box can
// reach.  A position lies in its cell's bounds except in the grid's edge
// cells, where cell_coord clamps, so only interior rows and cells are cut:
// a row whose y-range misses the box, and cells at either end whose x-range
// misses it.  What remains is one contiguous sorted range, still visited in
// canonical order; a candidate cut here lies outside the box, so every lane
// would reject it.  False when nothing of the row remains.
__device__ __forceinline__ bool row_span(const SceneP& P, int32_t gy, int32_t& xa, int32_t& xb, int32_t bx0,
                                         int32_t bx1, int32_t by0, int32_t by1) {
  const int32_t cs = (int32_t)P.cellShift, gw = (int32_t)P.gridW, gh = (int32_t)P.gridH;
  if (gy >= 1 && gy <= gh - 2 && (((gy + 1) << cs) <= by0 || (gy << cs) > by1)) return false;
  while (xa <= xb && xa >= 1 && xa <= gw - 2 && ((xa + 1) << cs) <= bx0) ++xa;
  while (xb >= xa && xb >= 1 && xb <= gw - 2 && (xb << cs) > bx1) --xb;
  return xa <= xb;
}

// Block boxes.  Per substep, after the canonical order: the bounding box of
// the live positions of every aligned 64-slot block of the sorted order, as
// (min x, min y, max x, max y); a block with no live slot has an empty box
// (min > max).  Heavy items of the sweep visit only the blocks whose boxes
// meet their reach box, found 64 blocks at a time with one vector load and
// a ballot: every candidate of a block passed over is out of reach of every
// lane the box covers, so only the work changes.  A corner pile-up of tens
// of thousands of slots on one point is then crossed in a few loads by the
// waves that cannot reach it, instead of one dependent load per 64 slots.
__device__ __forceinline__ bool box_miss(int4 b, int32_t x0, int32_t x1, int32_t y0, int32_t y1) {
  return b.x > x1 || b.z < x0 || b.y > y1 || b.w < y0;
}

__global__ __launch_bounds__(256) void k_block_box(Dev D) {
  const uint32_t q = slot_block(D) * 256u + threadIdx.x;
  const bool live = q < D.N && D.sidOf[q] != INACTIVE;
  const int2 p = live ? D.pred[q] : make_int2(0, 0);
  const int32_t x0 = wave_min(live ? p.x : I32_MAX), y0 = wave_min(live ? p.y : I32_MAX);
  const int32_t x1 = wave_max(live ? p.x : I32_MIN), y1 = wave_max(live ? p.y : I32_MIN);
  if ((threadIdx.x & 63u) == 0u) D.blkBox[q >> 6] = make_int4(x0, y0, x1, y1);
}

// Skipping blocks.  A sweep over a row range [q0, hi) keeps the mask of the
// aligned blocks of one 64-block group whose boxes meet its reach box [x0,
// x1] x [y0, y1], loaded with one vector load and a ballot.  skip_blocks
// moves q0 to the start of the first such block at or after it (hi when
// there is none): every slot passed over lies in a block whose box misses
// the reach box.  Uniform state and arguments; every lane active.
struct BlockMask {
  uint32_t group = 0xffffffffu;   // first block of the loaded group (none yet)
  uint64_t need = 0ull;
};
__device__ __forceinline__ uint32_t skip_blocks(const Dev& D, BlockMask& m, uint32_t q0, uint32_t hi, uint32_t lane,
                                                int32_t x0, int32_t x1, int32_t y0, int32_t y1) {
  const uint32_t b1 = (hi - 1u) >> 6;
  uint32_t b = q0 >> 6;
  while (b <= b1) {
    const uint32_t g = b & ~63u;
    if (m.group != g) {
      const uint32_t bk = g + lane;
      bool meet = false;
      if (bk <= b1) meet = !box_miss(D.blkBox[bk], x0, x1, y0, y1);
      m.group = g;
      m.need = __ballot(meet);
    }
    const uint64_t rest = m.need & (~0ull << (b - g));
    if (rest) return max(q0, (g + (uint32_t)__builtin_ctzll(rest)) << 6);
    b = g + 64u;
  }
  return hi;
}

__device__ __forceinline__ uint32_t morton_bin(const SceneP& P, uint32_t c0, int2 pr) {
  const uint32_t lc0 = c0 - P.cellBase;
  const int32_t x0 = (int32_t)((lc0 % P.gridW) << P.cellShift), y0 = (int32_t)((lc0 / P.gridW) << P.cellShift);
  const int32_t bs = maxi((int32_t)P.cellShift - 3, 0);
  const uint32_t bx = (uint32_t)clampi((pr.x - x0) >> bs, 0, 7), by = (uint32_t)clampi((pr.y - y0) >> bs, 0, 7);
  return (bx & 1u) | ((by & 1u) << 1) | ((bx & 2u) << 1) | ((by & 2u) << 2) | ((bx & 4u) << 2) | ((by & 4u) << 3);
}

// Exact in-range counts for the debug build.  Once every list of a wave is
// full, a further candidate changes only its lane's truncation count (and,
// in scenes with SEAL or CONGEAL particles, the seen bits), so in a bounded
// scene without such particles a one-cell chunk can take the rest of each
// lane's count at once: the number of live candidates not yet visited,
// other than itself, that the sweep's test accepts.  In a cell with a bin table
// (a packed cell, binned by k_heavy_order) a bin whose box lies wholly
// inside a lane's radius adds its live count to that lane and one wholly
// outside adds nothing; only bins that the radius crosses for some lane are
// tested slot by slot.  The test is monotone: r2q below is non-decreasing in
// each of |dx| and |dy|, so the box's farthest and nearest corners bound
// every slot in it.
__device__ __forceinline__ uint32_t r2q_abs(uint32_t a, uint32_t b) {
  // |d| < 2^29 in a bounded scene; past HS every pair is out of range
  const uint32_t ax = min((a + 8u) >> 4, (uint32_t)HS + 1u), ay = min((b + 8u) >> 4, (uint32_t)HS + 1u);
  return __umul24(ax, ax) + __umul24(ay, ay);
}

// Lanes with `on` count the slots map(i), i in [a, b), that are live, inside
// the chunk's reach box, not their own slot pk, and in range by the sweep's
// one-cell test (biased positions, |d| < 2^19 inside the reach box of a
// one-cell chunk).  Uniform arguments; every lane active.
__device__ __forceinline__ uint32_t count_exact(const Dev& D, uint4* stg, const uint32_t* perm, uint32_t a, uint32_t b,
                                                bool on, uint32_t pk, uint32_t pux, uint32_t puy, int32_t bx0,
                                                int32_t bx1, int32_t by0, int32_t by1, uint32_t lane) {
  uint32_t n = 0;
  for (uint32_t i0 = a; i0 < b; i0 += 64u) {
    const uint32_t i = i0 + lane;
    uint32_t q = 0, ci = INF_DEAD;
    int2 cp = make_int2(0, 0);
    if (i < b) {
      q = perm ? perm[i] : i;
      cp = D.pred[q];
      ci = D.info[q];
    }
    const bool want = !(ci & INF_DEAD) && cp.x >= bx0 && cp.x <= bx1 && cp.y >= by0 && cp.y <= by1;
    const uint64_t todo = __ballot(want);
    if (todo == 0ull) continue;
    // the block's candidates in reach, staged in the wave's LDS row and read
    // back by every lane (as in nbr_chunk_staged)
    if (want)
      stg[__builtin_amdgcn_mbcnt_hi((uint32_t)(todo >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)todo, 0u))] =
          make_uint4((uint32_t)cp.x ^ 0x80000000u, (uint32_t)cp.y ^ 0x80000000u, q, 0u);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t m = (uint32_t)__builtin_popcountll(todo);
#pragma unroll 4
    for (uint32_t j = 0; j < m; ++j) {
      const uint4 c = stg[j];
      const uint32_t ax = (max(pux, c.x) - min(pux, c.x) + 8u) >> 4;
      const uint32_t ay = (max(puy, c.y) - min(puy, c.y) + 8u) >> 4;
      n += (uint32_t)on & (uint32_t)(c.z != pk) & (uint32_t)(__umul24(ax, ax) + __umul24(ay, ay) < (uint32_t)H2S);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");   // reads done before the next block's stores
  }
  return n;
}

// A lane's in-range count over the whole of binned cell c (bin table at
// `slot`); c0 is the chunk's own cell and mybin the lane's bin in it.
__device__ __forceinline__ uint32_t bins_count(const Dev& D, uint4* stg, uint32_t slot, uint32_t lo, uint32_t c, uint32_t c0,
                                               uint32_t mybin, bool live, uint32_t pk, int2 pi, int32_t bx0,
                                               int32_t bx1, int32_t by0, int32_t by1, uint32_t lane) {
  const uint32_t pux = (uint32_t)pi.x ^ 0x80000000u, puy = (uint32_t)pi.y ^ 0x80000000u;
  const uint4* T = D.binTab + (size_t)slot * 64u;
  const int4* B = D.binBox + (size_t)slot * 64u;
  uint32_t n = 0;
  for (uint32_t b = 0; b < 64u; ++b) {
    const uint4 tb = T[b];
    const uint32_t nlive = (uint32_t)__builtin_amdgcn_readfirstlane((int)tb.z);
    if (nlive == 0u) continue;
    const int4 bb = B[b];
    const int32_t x0 = __builtin_amdgcn_readfirstlane(bb.x), y0 = __builtin_amdgcn_readfirstlane(bb.y);
    const int32_t x1 = __builtin_amdgcn_readfirstlane(bb.z), y1 = __builtin_amdgcn_readfirstlane(bb.w);
    if (x0 > bx1 || x1 < bx0 || y0 > by1 || y1 < by0) continue;   // out of every lane's reach
    const uint32_t fx = (uint32_t)maxi(abs(pi.x - x0), abs(pi.x - x1));
    const uint32_t fy = (uint32_t)maxi(abs(pi.y - y0), abs(pi.y - y1));
    const uint32_t nx = (uint32_t)(pi.x < x0 ? x0 - pi.x : (pi.x > x1 ? pi.x - x1 : 0));
    const uint32_t ny = (uint32_t)(pi.y < y0 ? y0 - pi.y : (pi.y > y1 ? pi.y - y1 : 0));
    const bool full = live && r2q_abs(fx, fy) < (uint32_t)H2S;
    const bool part = live && !full && r2q_abs(nx, ny) < (uint32_t)H2S;
    if (full) n += nlive - ((c == c0 && b == mybin) ? 1u : 0u);
    if (__ballot(part)) {
      const uint32_t first = (uint32_t)__builtin_amdgcn_readfirstlane((int)tb.x);
      const uint32_t end = (uint32_t)__builtin_amdgcn_readfirstlane((int)tb.y);
      n += count_exact(D, stg, D.hperm + lo, first, end, part, pk, pux, puy, bx0, bx1, by0, by1, lane);
    }
  }
  return n;
}

// A lane's in-range count over the slots [a, b) of stencil row gy, cells xlo
// .. xhi, of one-cell tile c0 (bounded scene, cell no wider than 2^18; 0 for
// a lane that is not live).  The chunk's reach box [bx0, bx1] x [by0, by1]
// holds every lane's reach.  A binned cell of more than BIN_MIN slots in the
// range is counted from its bins, less the slots of the cell outside the
// range counted exactly when those are fewer.
constexpr uint32_t BIN_MIN = 256;
__device__ __forceinline__ uint32_t range_count(const Dev& D, uint4* stg, const SceneP& P, int32_t gy, int32_t xlo, int32_t xhi,
                                                uint32_t a, uint32_t b, uint32_t c0, bool live, uint32_t pk,
                                                int2 pi, int32_t bx0, int32_t bx1, int32_t by0, int32_t by1,
                                                uint32_t lane) {
  const int32_t sh = P.posShift;
  const uint32_t pux = (uint32_t)pi.x ^ 0x80000000u, puy = (uint32_t)pi.y ^ 0x80000000u;
  const uint32_t rowc = P.cellBase + (uint32_t)gy * P.gridW;
  uint32_t n = 0;
  for (int32_t x = xlo; x <= xhi; ++x) {
    const uint32_t c = rowc + (uint32_t)x;
    const uint32_t lo = (uint32_t)((int32_t)D.cellStart[c] + sh);
    const uint32_t hi = (uint32_t)((int32_t)D.cellStart[c + 1] + sh);
    const uint32_t u = max(lo, a), v = min(hi, b);
    if (u >= v) continue;
    const uint32_t slot = (uint32_t)__builtin_amdgcn_readfirstlane((int)D.cellBin[c]);
    if (slot == BIN_NONE || v - u <= BIN_MIN || (u - lo) + (hi - v) >= v - u) {
      n += count_exact(D, stg, nullptr, u, v, live, pk, pux, puy, bx0, bx1, by0, by1, lane);
      continue;
    }
    n += bins_count(D, stg, slot, lo, c, c0, morton_bin(P, c0, pi), live, pk, pi, bx0, bx1, by0, by1, lane);
    n -= count_exact(D, stg, nullptr, lo, u, live, pk, pux, puy, bx0, bx1, by0, by1, lane);
    n -= count_exact(D, stg, nullptr, v, hi, live, pk, pux, puy, bx0, bx1, by0, by1, lane);
  }
  return n;
}

// The in-range count of the candidates a one-cell sweep has not yet visited:
// row gy0 from slot q0 on, then every later row of the stencil (rows below
// gy0 + 1 .. cy + 1, cells xlo .. xhi).
__device__ __forceinline__ uint32_t rest_count(const Dev& D, uint4* stg, const SceneP& P, int32_t cy, int32_t xlo, int32_t xhi,
                                               int32_t gy0, uint32_t q0, uint32_t c0, bool live, uint32_t pk,
                                               int2 pi, int32_t bx0, int32_t bx1, int32_t by0, int32_t by1,
                                               uint32_t lane) {
  uint32_t n = range_count(D, stg, P, gy0, xlo, xhi, q0, 0xffffffffu, c0, live, pk, pi, bx0, bx1, by0, by1, lane);
  for (int32_t gy = gy0 + 1; gy <= mini(cy + 1, (int32_t)P.gridH - 1); ++gy)
    n += range_count(D, stg, P, gy, xlo, xhi, 0u, 0xffffffffu, c0, live, pk, pi, bx0, bx1, by0, by1, lane);
  return n;
}

// The sweep's tail for one particle (lane L of the wave) whose list is not
// yet full while the rest of the wave is: the remaining candidates, from row
// dy0 at slot q0s on, are tested one per lane, 64 at a time, and the
// accepted ones taken in canonical order (ballot rank) until the list is
// full.  Its iteration-0 solveA terms are summed over the lanes and added to
// lane L's sums (wrapping u32 sums, so the grouping changes nothing).  Used
// only without debug counters and without SEAL/CONGEAL particles in the
// scene, where a full list ends a particle's sweep; then a cell that walls
// have packed with thousands of particles costs a few lanes' worth of work
// per chunk instead of one broadcast per candidate.
// The tail costs about 25 instructions per 64 candidates per open list, the
// broadcast sweep about 20 per candidate for the whole wave, so up to about
// 50 open lists the tail is cheaper; 32 keeps a margin.
constexpr uint32_t TAIL_LANES = 32;   // at most this many unfilled lists switch to the tail
constexpr uint32_t DUP_MIN = 16;      // equal neighbours in a block that select the grouped loop

template <bool BND, bool WIN>
__device__ __forceinline__ void nbr_tail(const Dev& D, const SceneP& P, int32_t cy, int32_t xlo, int32_t xhi,
                                         int32_t dy0, uint32_t q0s, uint32_t lane, uint32_t L, int2 pi,
                                         uint32_t pk, uint32_t inf, int32_t cxk, uint32_t& cnt, SolveAAcc& acc,
                                         Dbg& g) {
  const int32_t sh = P.posShift;
  const int2 pL = make_int2(__builtin_amdgcn_readlane(pi.x, L), __builtin_amdgcn_readlane(pi.y, L));
  const uint32_t pkL = (uint32_t)__builtin_amdgcn_readlane((int)pk, L);
  const bool fluidL = ((uint32_t)__builtin_amdgcn_readlane((int)inf, L) & INF_FLUID) != 0u;
  const int32_t cxL = WIN ? __builtin_amdgcn_readlane(cxk, L) : 0;
  uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)cnt, L);
  SolveAAcc part{0, 0, 0, 0};
  for (int32_t dy = dy0; dy <= 1 && c < MAXNBR; ++dy) {
    const int32_t gy = cy + dy;
    if (gy < 0 || gy >= (int32_t)P.gridH) continue;
    int32_t xa = xlo, xb = xhi;
    // the particle's own reach box (bounded scenes; positions within +-2^28)
    if (BND && !row_span(P, gy, xa, xb, pL.x - REACH, pL.x + REACH, pL.y - REACH, pL.y + REACH)) continue;
    const uint32_t rowc = P.cellBase + (uint32_t)gy * P.gridW;
    const uint32_t rlo = (uint32_t)((int32_t)D.cellStart[rowc + (uint32_t)xa] + sh);
    const uint32_t rhi = (uint32_t)((int32_t)D.cellStart[rowc + (uint32_t)xb + 1u] + sh);
    BlockMask bm;
    for (uint32_t q0 = dy == dy0 ? max(q0s, rlo) : rlo; q0 < rhi && c < MAXNBR; q0 += 64u) {
      if constexpr (BND) {
        // bounded scenes: positions within +-2^28, so the box cannot wrap
        q0 = skip_blocks(D, bm, q0, rhi, lane, pL.x - REACH, pL.x + REACH, pL.y - REACH, pL.y + REACH);
        if (q0 >= rhi) break;
      }
      const uint32_t q = q0 + lane;
      bool ok = q < rhi;
      int2 cp = make_int2(0, 0);
      uint32_t ci = 0;
      int32_t cc = 0;
      if (ok) {
        cp = D.pred[q];
        ci = D.info[q];
        if (WIN) cc = (int32_t)(D.cellY[q] - rowc);
      }
      ok = ok && !(ci & INF_DEAD) && q != pkL;
      if (WIN) ok = ok && cc >= cxL - 1 && cc <= cxL + 1;
      // WIN false: one cell of <= 2^18 ticks, so |d| < 2^18 + REACH < 2^19
      // for every candidate of its rows that can be accepted
      const int2 d = pdelta<BND>(pL, cp, g);
      const int32_t idx = ok ? ((BND && !WIN) ? lut_idx_near(d) : lut_idx<BND>(d, g)) : -1;
      ok = ok && idx >= 0;
      const uint64_t m = __ballot(ok);
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      const uint32_t room = MAXNBR - c;
      if (ok && rank < room) {
        D.nbr[(size_t)(c + rank) * D.NS + pkL] = q | ((ci & INF_FLUID) ? 0u : NBR_BND);
        if (fluidL) solveA_term<BND>(part, d, D.lut4[idx], g);
      }
      c += min((uint32_t)__popcll(m), room);
    }
  }
  part.rho = (int32_t)wave_sum((uint32_t)part.rho);
  part.S = (int32_t)wave_sum((uint32_t)part.S);
  part.sgx = (int32_t)wave_sum((uint32_t)part.sgx);
  part.sgy = (int32_t)wave_sum((uint32_t)part.sgy);
  if (lane == L) {
    cnt = c;
    acc.rho = wadd(acc.rho, part.rho);
    acc.S = wadd(acc.S, part.S);
    acc.sgx = wadd(acc.sgx, part.sgx);
    acc.sgy = wadd(acc.sgy, part.sgy);
  }
}

template <bool BND, bool WIN, int PPL, bool STRIP>
__device__ __forceinline__ void nbr_chunk(const Dev& D, uint4* stg, uint32_t c0, uint32_t ncell, uint32_t s,
                                          uint32_t plo, uint32_t phi, uint32_t lane, int32_t W0, Dbg& g,
                                          const uint32_t* perm) {
  const SceneP& P = D.sp[s];
  const int32_t sh = P.posShift;
  const uint32_t gridW = P.gridW, gridH = P.gridH, cellBase = P.cellBase;
  const uint32_t lc0 = c0 - cellBase;
  const int32_t cy = (int32_t)(lc0 / gridW), cx0 = (int32_t)(lc0 % gridW);
  const int32_t xlo = maxi(cx0 - 1, 0), xhi = mini(cx0 + (int32_t)ncell, (int32_t)gridW - 1);
  const uint32_t myRow = cellBase + (uint32_t)cy * gridW;
  // without debug counters, and with no SEAL/CONGEAL particle in the scene,
  // a wave whose lists are all full has nothing left to find
  const bool stopFull = STRIP && D.sd[s].flagged == 0u;
  // with debug counters, and with no SEAL/CONGEAL particle in the scene, a
  // wave whose lists are all full takes the rest of each lane's count from
  // rest_count
  const bool binCount = !STRIP && D.dbgOn && D.sd[s].flagged == 0u;
  for (uint32_t pbase = plo; pbase < phi; pbase += 64u * PPL) {
    uint32_t pk[PPL], inf[PPL], cnt[PPL], bits[PPL], dropped[PPL];
    int2 pik[PPL];
    int32_t cxk[PPL];
    bool live[PPL], vld[PPL];
    SolveAAcc acc[PPL];
    int32_t mnx = I32_MAX, mny = I32_MAX, mxx = I32_MIN, mxy = I32_MIN;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      const uint32_t li = pbase + (uint32_t)k * 64u + lane;
      const bool valid = li < phi;
      vld[k] = valid;
      pk[k] = (perm && valid) ? perm[li] : li;
      inf[k] = valid ? D.info[pk[k]] : 0u;
      pik[k] = valid ? D.pred[pk[k]] : make_int2(0, 0);
      cxk[k] = valid ? (int32_t)(D.cellY[pk[k]] - myRow) : 0;
      live[k] = valid && !(inf[k] & INF_DEAD);
      cnt[k] = 0;
      bits[k] = 0;
      dropped[k] = 0;
      acc[k] = SolveAAcc{W0, 0, 0, 0};
      if (live[k]) {
        mnx = mini(mnx, pik[k].x);
        mxx = maxi(mxx, pik[k].x);
        mny = mini(mny, pik[k].y);
        mxy = maxi(mxy, pik[k].y);
      }
    }
    // the chunk's reach box (bounded scenes: no overflow within +-2^28 + REACH)
    int32_t bx0 = 0, bx1 = 0, by0 = 0, by1 = 0;
    if constexpr (BND) {
      bx0 = wave_min(mnx) - REACH;
      bx1 = wave_max(mxx) + REACH;
      by0 = wave_min(mny) - REACH;
      by1 = wave_max(mxy) + REACH;
    }
    bool done = false;
    for (int32_t dy = -1; dy <= 1 && !done; ++dy) {
      const int32_t gy = cy + dy;
      if (gy < 0 || gy >= (int32_t)gridH) continue;
      int32_t xa = xlo, xb = xhi;
      if (BND && !row_span(P, gy, xa, xb, bx0, bx1, by0, by1)) continue;
      const uint32_t rowc = cellBase + (uint32_t)gy * gridW;
      const uint32_t rlo = (uint32_t)((int32_t)D.cellStart[rowc + (uint32_t)xa] + sh);
      const uint32_t rhi = (uint32_t)((int32_t)D.cellStart[rowc + (uint32_t)xb + 1u] + sh);
      for (uint32_t q0 = rlo; q0 < rhi; q0 += 64u) {
        if constexpr (!STRIP && BND && !WIN && PPL == 1) {
          if (binCount && __all(!live[0] || cnt[0] >= MAXNBR)) {
            dropped[0] += rest_count(D, stg, P, cy, xlo, xhi, gy, q0, c0, live[0], pk[0], pik[0], bx0, bx1, by0, by1, lane);
            done = true;
            break;
          }
        }
        uint32_t open = 0;   // lists still filling, over the wave
        if (stopFull) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) open += (uint32_t)__popcll(__ballot(live[k] && cnt[k] < MAXNBR));
          if (open == 0u) {
            done = true;
            break;
          }
        }
        const uint32_t qn = (uint32_t)mini((int32_t)(rhi - q0), 64);
        const uint32_t ql = q0 + lane;
        int2 cp = make_int2(0, 0);
        uint32_t ci = 0;
        int32_t cc = 0;
        if (lane < qn) {
          cp = D.pred[ql];
          ci = D.info[ql];
          cc = (int32_t)(D.cellY[ql] - rowc);
        }
        // wait for this block's candidates here, so that the candidate
        // loop below never waits on the neighbour-list stores it issues
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        // a block mostly at one exact position (a corner pile-up): taken
        // position by position below, see nbr_chunk_staged
        bool grouped = false;
        if constexpr (BND && !WIN && PPL == 1) {
          const int32_t upx = __shfl_up(cp.x, 1), upy = __shfl_up(cp.y, 1);
          grouped = __builtin_popcountll(__ballot(lane > 0u && lane < qn && cp.x == upx && cp.y == upy)) >= DUP_MIN;
        }
        if (stopFull && !grouped && open <= TAIL_LANES) {
          // few lists left: finish each alone over the remaining candidates
#pragma unroll
          for (int k = 0; k < PPL; ++k) {
            uint64_t m = __ballot(live[k] && cnt[k] < MAXNBR);
            while (m) {
              const uint32_t L = (uint32_t)__builtin_ctzll(m);
              m &= m - 1ull;
              nbr_tail<BND, WIN>(D, P, cy, xlo, xhi, dy, q0, lane, L, pik[k], pk[k], inf[k], cxk[k], cnt[k],
                                 acc[k], g);
            }
          }
          done = true;
          break;
        }
        bool want = lane < qn && !(ci & INF_DEAD);   // nobody sees a corpse
        if constexpr (BND) want = want && cp.x >= bx0 && cp.x <= bx1 && cp.y >= by0 && cp.y <= by1;
        uint64_t todo = __ballot(want);
        if constexpr (BND && !WIN && PPL == 1) {
          if (grouped) {
            // as in nbr_chunk_staged: candidate j with every later todo
            // candidate of the block at its position, while each lane that
            // accepts them holds a full list (only `dropped` and the seen
            // bits change); otherwise j alone
            const uint64_t mS = __ballot(lane < qn && (ci & INF_SEAL));
            const uint64_t mC = __ballot(lane < qn && (ci & INF_CONGEAL));
            if (!D.dbgOn && __all(!live[0] || cnt[0] >= MAXNBR)) todo &= mS | mC;
            const uint32_t selfj = pk[0] - q0;
            const uint64_t mine = selfj < 64u ? (1ull << selfj) : 0ull;
            while (todo) {
              const uint32_t j = (uint32_t)__builtin_ctzll(todo);
              const int2 cj = make_int2(__builtin_amdgcn_readlane(cp.x, j), __builtin_amdgcn_readlane(cp.y, j));
              const uint64_t eq = __ballot(cp.x == cj.x && cp.y == cj.y) & todo;
              const int2 d = pdelta<true>(pik[0], cj, g);
              const int32_t idx = lut_idx_near(d);
              const uint64_t take = eq & ~mine;
              const bool accAll = live[0] && idx >= 0 && take != 0ull;
              if (__all(!accAll || cnt[0] >= MAXNBR)) {
                const uint64_t tk = accAll ? take : 0ull;
                bits[0] |= ((mS & tk) ? INF_SEAL : 0u) | ((mC & tk) ? INF_CONGEAL : 0u);
                dropped[0] += (uint32_t)__builtin_popcountll(tk);
                todo &= ~eq;
              } else {
                todo ^= 1ull << j;
                const uint32_t q = q0 + j;
                const uint32_t qi = (uint32_t)__builtin_amdgcn_readlane((int)ci, j);
                if (live[0] && q != pk[0] && idx >= 0) {
                  bits[0] |= qi & (INF_SEAL | INF_CONGEAL);
                  if (cnt[0] >= MAXNBR) {
                    dropped[0]++;
                  } else {
                    D.nbr[(size_t)cnt[0] * D.NS + pk[0]] = q | ((qi & INF_FLUID) ? 0u : NBR_BND);
                    cnt[0]++;
                    if (inf[0] & INF_FLUID) solveA_term<BND>(acc[0], d, D.lut4[idx], g);
                  }
                }
              }
            }
            continue;
          }
        }
        bool full = true;
#pragma unroll
        for (int k = 0; k < PPL; ++k) full = full && (!live[k] || cnt[k] >= MAXNBR);
        if (__all(full)) {
          // every list of the wave is full: a remaining candidate only adds
          // to `dropped` (debug counters) and to nbrN's SEAL / CONGEAL seen
          // bits, so without debug counters only candidates carrying those
          // bits are left (with none in the scene the sweep has stopped)
          if (!D.dbgOn) todo &= __ballot(lane < qn && (ci & (INF_SEAL | INF_CONGEAL)));
          if constexpr (BND) {
            // branch-free count; the bounded pair geometry counts no event,
            // and `live` gates every lane whose inputs are not a real pair
            const uint64_t mS = __ballot(lane < qn && (ci & INF_SEAL));
            const uint64_t mC = __ballot(lane < qn && (ci & INF_CONGEAL));
            while (todo) {
              const uint32_t j = (uint32_t)__builtin_ctzll(todo);
              todo &= todo - 1ull;
              const int2 cj = make_int2(__builtin_amdgcn_readlane(cp.x, j), __builtin_amdgcn_readlane(cp.y, j));
              const int32_t qc = WIN ? __builtin_amdgcn_readlane(cc, j) : 0;
              const uint32_t q = q0 + j;
              const uint32_t fb = (((mS >> j) & 1ull) ? INF_SEAL : 0u) | (((mC >> j) & 1ull) ? INF_CONGEAL : 0u);
#pragma unroll
              for (int k = 0; k < PPL; ++k) {
                const int2 d = pdelta<true>(pik[k], cj, g);
                const int32_t idx = WIN ? lut_idx<true>(d, g) : lut_idx_near(d);
                bool a = live[k] && q != pk[k] && idx >= 0;
                if (WIN) a = a && qc >= cxk[k] - 1 && qc <= cxk[k] + 1;
                dropped[k] += a ? 1u : 0u;
                bits[k] |= a ? fb : 0u;
              }
            }
            continue;
          }
        }
        while (todo) {
          const uint32_t j = (uint32_t)__builtin_ctzll(todo);
          todo &= todo - 1ull;
          const int32_t px = __builtin_amdgcn_readlane(cp.x, j);
          const int32_t py = __builtin_amdgcn_readlane(cp.y, j);
          const uint32_t qi = (uint32_t)__builtin_amdgcn_readlane((int)ci, j);
          const int32_t qc = WIN ? __builtin_amdgcn_readlane(cc, j) : 0;
          const uint32_t q = q0 + j;
#pragma unroll
          for (int k = 0; k < PPL; ++k) {
            if (!live[k] || q == pk[k]) continue;
            if (WIN && (qc < cxk[k] - 1 || qc > cxk[k] + 1)) continue;
            const int2 d = pdelta<BND>(pik[k], make_int2(px, py), g);
            // WIN false: one cell of <= 2^18 ticks, so |d| < 2^18 + REACH < 2^19
            const int32_t idx = (BND && !WIN) ? lut_idx_near(d) : lut_idx<BND>(d, g);
            if (idx < 0) continue;
            bits[k] |= qi & (INF_SEAL | INF_CONGEAL);
            if (cnt[k] >= MAXNBR) {
              dropped[k]++;
              continue;
            }
            D.nbr[(size_t)cnt[k] * D.NS + pk[k]] = q | ((qi & INF_FLUID) ? 0u : NBR_BND);
            cnt[k]++;
            if (inf[k] & INF_FLUID) solveA_term<BND>(acc[k], d, D.lut4[idx], g);
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      if (!vld[k]) continue;
      D.nbrN[pk[k]] = live[k] ? (cnt[k] | ((bits[k] & INF_SEAL) ? NBR_SEAL_BIT : 0u) |
                                 ((bits[k] & INF_CONGEAL) ? NBR_CONGEAL_BIT : 0u))
                              : 0u;
      D.drop[pk[k]] = D.dbgOn ? dropped[k] : 0u;
      if (inf[k] & INF_FLUID) solveA_finish(D, pk[k], pik[k], acc[k], MAT(D, (uint32_t)D.ptY[pk[k]].x)[M_EPS], g);
    }
  }
}

// The one-wave form of the sweep (one particle per lane) with the lists
// staged in LDS.  The candidate loop only tests and records: an accepted
// candidate is stored as 16 bits, (boundary << 15) | its offset in the
// three row ranges laid end to end, at list[k][lane].  Afterwards the lists
// are written out transposed, one coalesced 256-byte store per k, and
// iteration 0 of solveA runs over them with every lane busy, instead of
// inside the candidate loop where only the few accepting lanes of each
// candidate would run it.  Returns false, having done nothing, when the rows
// hold more candidates than the 15-bit offset reaches; the caller then takes
// nbr_chunk.  (A corner pile-up of a 256 x 128 wu box holds up to about
// 27,000.)
constexpr uint32_t CAND_OFF_MAX = 1u << 15;

template <bool BND, bool WIN, bool STRIP>
__device__ __forceinline__ bool nbr_chunk_staged(const Dev& D, uint16_t* lst, uint4* stg, uint32_t c0,
                                                 uint32_t ncell, uint32_t s, uint32_t plo, uint32_t phi,
                                                 uint32_t lane, int32_t W0, Dbg& g, const uint32_t* perm) {
  const SceneP& P = D.sp[s];
  const int32_t sh = P.posShift;
  const uint32_t gridW = P.gridW, gridH = P.gridH, cellBase = P.cellBase;
  const uint32_t lc0 = c0 - cellBase;
  const int32_t cy = (int32_t)(lc0 / gridW), cx0 = (int32_t)(lc0 % gridW);
  const int32_t xlo = maxi(cx0 - 1, 0), xhi = mini(cx0 + (int32_t)ncell, (int32_t)gridW - 1);
  const uint32_t myRow = cellBase + (uint32_t)cy * gridW;
  const bool valid = plo + lane < phi;
  const uint32_t p = (perm && valid) ? perm[plo + lane] : plo + lane;
  const uint32_t inf = valid ? D.info[p] : 0u;
  const int2 pi = valid ? D.pred[p] : make_int2(0, 0);
  const int32_t cxk = valid ? (int32_t)(D.cellY[p] - myRow) : 0;
  const bool live = valid && !(inf & INF_DEAD);
  uint32_t cnt = 0, bits = 0, dropped = 0, nin = 0;   // nin: in-range candidates taken or dropped
  int32_t bx0 = 0, bx1 = 0, by0 = 0, by1 = 0;
  if constexpr (BND) {
    bx0 = wave_min(live ? pi.x : I32_MAX) - REACH;
    bx1 = wave_max(live ? pi.x : I32_MIN) + REACH;
    by0 = wave_min(live ? pi.y : I32_MAX) - REACH;
    by1 = wave_max(live ? pi.y : I32_MIN) + REACH;
  }
  // each row's candidates, cut to the reach box (row_span)
  uint32_t rlo[3], rhi[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int32_t gy = cy + r - 1;
    rlo[r] = rhi[r] = 0;
    if (gy < 0 || gy >= (int32_t)gridH) continue;
    int32_t xa = xlo, xb = xhi;
    if (BND && !row_span(P, gy, xa, xb, bx0, bx1, by0, by1)) continue;
    const uint32_t rowc = cellBase + (uint32_t)gy * gridW;
    rlo[r] = (uint32_t)((int32_t)D.cellStart[rowc + (uint32_t)xa] + sh);
    rhi[r] = (uint32_t)((int32_t)D.cellStart[rowc + (uint32_t)xb + 1u] + sh);
  }
  // the rows' starts in the end-to-end offset space of the list entries
  const uint32_t cb1 = rhi[0] - rlo[0], cb2 = cb1 + (rhi[1] - rlo[1]);
  if (cb2 + (rhi[2] - rlo[2]) > CAND_OFF_MAX) return false;
  const uint32_t cbase[3] = {0u, cb1, cb2};
  const bool stopFull = STRIP && D.sd[s].flagged == 0u;   // as in nbr_chunk
  const bool binCount = !STRIP && D.dbgOn && D.sd[s].flagged == 0u;   // as in nbr_chunk
  // The candidate loops below are branch-free per candidate: a lane counts
  // its in-range candidates in nin (its list keeps the first MAXNBR, so cnt =
  // min(nin, MAXNBR) and the truncation count is nin - cnt), and every lane
  // stores the candidate's entry at its own next slot, lst[min(nin,
  // MAXNBR)][lane], whether or not it takes the candidate: a slot past the
  // lane's count is never read and the next accepted candidate overwrites
  // it, and a full list writes the spare row MAXNBR.  Per candidate the wave then issues a few
  // scalar instructions instead of a branch per condition (the scalar unit,
  // shared by the CU's SIMDs, was the limit).  In the one-cell form (!WIN)
  // the radius test takes |a - b| + 8 of positions biased by 2^31 (unsigned
  // order equals signed order, and |d| < 2^19), one v_sad_u32 per axis.
  const uint32_t pux = (uint32_t)pi.x ^ 0x80000000u, puy = (uint32_t)pi.y ^ 0x80000000u;
  bool done = false;
#pragma unroll
  for (uint32_t r = 0; r < 3; ++r) {
    if (done) break;
    const uint32_t rowc = cellBase + (uint32_t)(cy + (int32_t)r - 1) * gridW;
    // each block's candidates are loaded while the block before it is tested
    int2 ncp = make_int2(0, 0);
    uint32_t nci = 0;
    int32_t ncc = 0;
    if (rlo[r] + lane < rhi[r]) {
      ncp = D.pred[rlo[r] + lane];
      nci = D.info[rlo[r] + lane];
      if constexpr (WIN) ncc = (int32_t)(D.cellY[rlo[r] + lane] - rowc);
    }
    for (uint32_t q0 = rlo[r]; q0 < rhi[r]; q0 += 64u) {
      if (stopFull && __all(!live || cnt >= MAXNBR)) {
        done = true;
        break;
      }
      if constexpr (!STRIP && BND && !WIN) {
        if (binCount && __all(!live || cnt >= MAXNBR)) {
          // every list is full: the rest of each lane's in-range count at once
          nin += rest_count(D, stg, P, cy, xlo, xhi, cy + (int32_t)r - 1, q0, c0, live, p, pi, bx0, bx1, by0, by1, lane);
          done = true;
          break;
        }
      }
      const uint32_t qn = (uint32_t)mini((int32_t)(rhi[r] - q0), 64);
      const int2 cp = ncp;
      const uint32_t ci = nci;
      const int32_t cc = ncc;
      ncp = make_int2(0, 0);
      nci = 0;
      ncc = 0;
      if (q0 + 64u + lane < rhi[r]) {
        ncp = D.pred[q0 + 64u + lane];
        nci = D.info[q0 + 64u + lane];
        if constexpr (WIN) ncc = (int32_t)(D.cellY[q0 + 64u + lane] - rowc);
      }
      bool want = lane < qn && !(ci & INF_DEAD);   // nobody sees a corpse
      if constexpr (BND) want = want && cp.x >= bx0 && cp.x <= bx1 && cp.y >= by0 && cp.y <= by1;
      uint64_t todo = __ballot(want);
      // the candidates' flag bits as wave masks: read per candidate with
      // scalar shifts instead of a vector lane read
      const uint64_t mBnd = __ballot(lane < qn && !(ci & INF_FLUID));
      const uint64_t mSeal = __ballot(lane < qn && (ci & INF_SEAL));
      const uint64_t mCong = __ballot(lane < qn && (ci & INF_CONGEAL));
      const bool flagged = (mSeal | mCong) != 0ull;   // uniform: no SEAL/CONGEAL candidate in this block
      const uint32_t selfj = p - q0;                 // lane index of the particle itself, if in this block
      const uint32_t cux = (uint32_t)cp.x ^ 0x80000000u, cuy = (uint32_t)cp.y ^ 0x80000000u;
      const bool allFull = __all(!live || cnt >= MAXNBR);
      // every list full: only `dropped` (debug counters) and the SEAL /
      // CONGEAL seen bits can change, so without debug counters only
      // candidates carrying those bits are left
      if (allFull && !D.dbgOn) todo &= mSeal | mCong;
      // this lane's candidate as a list entry, read per candidate with one lane read
      const uint32_t ent = cbase[r] + q0 - rlo[r] + lane + (((mBnd >> lane) & 1ull) ? CAND_OFF_MAX : 0u);
      bool grouped = false;
      if constexpr (BND && !WIN) {
        // A block mostly at one exact position: wall_clamp stacks a pile-up
        // on a corner point, thousands of slots deep.  Lanes that cannot
        // reach it never fill, so without this every such candidate costs an
        // iteration on the pile's longest wave.
        const uint32_t upx = (uint32_t)__shfl_up((int)cux, 1), upy = (uint32_t)__shfl_up((int)cuy, 1);
        grouped = __builtin_popcountll(__ballot(lane > 0u && lane < qn && cux == upx && cuy == upy)) >= DUP_MIN;
      }
      if constexpr (BND && !WIN) {
        if (!grouped && !flagged) {
          // The block's todo candidates in order, staged in this wave's LDS
          // row as {x, y, list entry, slot} and read back by every lane: one
          // LDS broadcast per candidate instead of a scalar pick and lane
          // reads.  Each lane still takes them in canonical order.
          if ((todo >> lane) & 1ull)
            stg[__builtin_amdgcn_mbcnt_hi((uint32_t)(todo >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)todo, 0u))] =
                make_uint4(cux, cuy, ent, q0 + lane);
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          const uint32_t n = (uint32_t)__builtin_popcountll(todo);
          if (allFull) {
            // only the truncation count changes (debug counters)
#pragma unroll 4
            for (uint32_t j = 0; j < n; ++j) {
              const uint4 c = stg[j];
              const uint32_t ax = (max(pux, c.x) - min(pux, c.x) + 8u) >> 4;
              const uint32_t ay = (max(puy, c.y) - min(puy, c.y) + 8u) >> 4;
              nin += (uint32_t)live & (uint32_t)(c.w != p) & (uint32_t)(__umul24(ax, ax) + __umul24(ay, ay) < (uint32_t)H2S);
            }
          } else {
#pragma unroll 4
            for (uint32_t j = 0; j < n; ++j) {
              const uint4 c = stg[j];
              const uint32_t ax = (max(pux, c.x) - min(pux, c.x) + 8u) >> 4;
              const uint32_t ay = (max(puy, c.y) - min(puy, c.y) + 8u) >> 4;
              lst[min(nin, MAXNBR) * 64u + lane] = (uint16_t)c.z;
              nin += (uint32_t)live & (uint32_t)(c.w != p) & (uint32_t)(__umul24(ax, ax) + __umul24(ay, ay) < (uint32_t)H2S);
            }
          }
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");   // reads done before the next block's stores
          cnt = min(nin, MAXNBR);
          continue;
        }
      }
      if (grouped) {
        // Candidate j is taken together with every later todo candidate of
        // the block at the same position, whenever each lane that accepts
        // them already holds MAXNBR entries: the list is then unchanged and
        // the count and seen bits are order-free.  Otherwise j alone, as below.
        const uint64_t mine = selfj < 64u ? (1ull << selfj) : 0ull;
        while (todo) {
          const uint32_t j = (uint32_t)__builtin_ctzll(todo);
          const uint32_t jx = (uint32_t)__builtin_amdgcn_readlane((int)cux, j);
          const uint32_t jy = (uint32_t)__builtin_amdgcn_readlane((int)cuy, j);
          const uint64_t eq = __ballot(cux == jx && cuy == jy) & todo;
          const uint32_t ax = (max(pux, jx) - min(pux, jx) + 8u) >> 4;
          const uint32_t ay = (max(puy, jy) - min(puy, jy) + 8u) >> 4;
          const bool inr = __umul24(ax, ax) + __umul24(ay, ay) < (uint32_t)H2S;
          const uint64_t take = eq & ~mine;
          const bool accAll = live && inr && take != 0ull;
          if (__all(!accAll || nin >= MAXNBR)) {
            if (flagged)
              bits |= accAll ? (((mSeal & take) ? INF_SEAL : 0u) | ((mCong & take) ? INF_CONGEAL : 0u)) : 0u;
            nin += accAll ? (uint32_t)__builtin_popcountll(take) : 0u;
            todo &= ~eq;
          } else {
            todo ^= 1ull << j;
            const bool acc = live && j != selfj && inr;
            if (flagged)
              bits |= acc ? ((((mSeal >> j) & 1ull) ? INF_SEAL : 0u) | (((mCong >> j) & 1ull) ? INF_CONGEAL : 0u)) : 0u;
            if (!allFull) lst[min(nin, MAXNBR) * 64u + lane] = (uint16_t)__builtin_amdgcn_readlane((int)ent, j);
            nin += acc ? 1u : 0u;
          }
        }
      }
      while (todo) {
        const uint32_t j = (uint32_t)__builtin_ctzll(todo);
        todo ^= 1ull << j;
        bool inr;
        if constexpr (BND && !WIN) {
          const uint32_t jx = (uint32_t)__builtin_amdgcn_readlane((int)cux, j);
          const uint32_t jy = (uint32_t)__builtin_amdgcn_readlane((int)cuy, j);
          const uint32_t ax = (max(pux, jx) - min(pux, jx) + 8u) >> 4;
          const uint32_t ay = (max(puy, jy) - min(puy, jy) + 8u) >> 4;
          inr = __umul24(ax, ax) + __umul24(ay, ay) < (uint32_t)H2S;
        } else {
          const int2 cj = make_int2(__builtin_amdgcn_readlane(cp.x, j), __builtin_amdgcn_readlane(cp.y, j));
          inr = lut_idx<BND>(pdelta<BND>(pi, cj, g), g) >= 0;
        }
        bool acc = live && j != selfj && inr;
        if constexpr (WIN) {
          const int32_t qc = __builtin_amdgcn_readlane(cc, j);
          acc = acc && qc >= cxk - 1 && qc <= cxk + 1;
        }
        if (flagged)
          bits |= acc ? ((((mSeal >> j) & 1ull) ? INF_SEAL : 0u) | (((mCong >> j) & 1ull) ? INF_CONGEAL : 0u)) : 0u;
        if (!allFull) lst[min(nin, MAXNBR) * 64u + lane] = (uint16_t)__builtin_amdgcn_readlane((int)ent, j);
        nin += acc ? 1u : 0u;
      }
      cnt = min(nin, MAXNBR);
    }
  }
  dropped = nin - cnt;
  // write-out (coalesced per k) and iteration 0 of solveA over the lists,
  // NBR_UNROLL entries at a time with their gathers issued together
  SolveAAcc acc{W0, 0, 0, 0};
  const uint32_t kmax = (uint32_t)wave_max((int32_t)cnt);
  const bool fl = live && (inf & INF_FLUID);
  for (uint32_t k = 0; k < kmax; k += NBR_UNROLL) {
    uint32_t q[NBR_UNROLL];
    int2 d[NBR_UNROLL];
    int32_t idx[NBR_UNROLL];
    int4 L[NBR_UNROLL];
#pragma unroll
    for (uint32_t u = 0; u < NBR_UNROLL; ++u) {
      q[u] = p;
      if (k + u < cnt) {
        const uint32_t e = lst[(k + u) * 64u + lane];
        const uint32_t o = e & (CAND_OFF_MAX - 1u);
        q[u] = o < cb1 ? rlo[0] + o : (o < cb2 ? rlo[1] + (o - cb1) : rlo[2] + (o - cb2));
        D.nbr[(size_t)(k + u) * D.NS + p] = q[u] | ((e & CAND_OFF_MAX) ? NBR_BND : 0u);
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < NBR_UNROLL; ++u) {
      d[u] = pdelta<BND>(pi, valid ? D.pred[q[u]] : pi, g);
      idx[u] = (fl && k + u < cnt) ? ((BND && !WIN) ? lut_idx_near(d[u]) : lut_idx<BND>(d[u], g)) : -1;
      L[u] = D.lut4[maxi(idx[u], 0)];
    }
#pragma unroll
    for (uint32_t u = 0; u < NBR_UNROLL; ++u)
      if (idx[u] >= 0) solveA_term<BND>(acc, d[u], L[u], g);
  }
  if (valid) {
    D.nbrN[p] = live ? (cnt | ((bits & INF_SEAL) ? NBR_SEAL_BIT : 0u) | ((bits & INF_CONGEAL) ? NBR_CONGEAL_BIT : 0u))
                     : 0u;
    D.drop[p] = D.dbgOn ? dropped : 0u;
    if (inf & INF_FLUID) solveA_finish(D, p, pi, acc, MAT(D, (uint32_t)D.ptY[p].x)[M_EPS], g);
  }
  return true;
}

// Spatial chunks for packed cells.  Which particles share a wave of the
// sweep changes no result (each builds its own list in its own canonical
// order, and every sum is order-free), so a one-cell tile of the heavy list
// with more than one chunk is chunked by position: k_heavy_order sorts its
// slots into 8 x 8 Morton bins of the cell (a counting sort in LDS, one
// workgroup per tile, bin order within a bin left to the LDS atomics) and
// writes them to hperm.  A chunk of 64 neighbouring particles then has a
// tight reach box, so whole blocks of far candidates drop out of its ballot.
// The same pass records, per substep, each bin of such a cell: its range in
// hperm, its live (not DEAD) slots and the box of their predicted positions
// (binTab / binBox at the cell's cellBin slot), for bins_count.
__global__ __launch_bounds__(256) void k_heavy_order(Dev D) {
  __shared__ uint32_t bins[64], bstart[64], blive[64];
  __shared__ int32_t bx0[64], by0[64], bx1[64], by1[64];
  __shared__ uint32_t slotS;
  const uint32_t n = (uint32_t)__builtin_amdgcn_readfirstlane((int)D.listCount[6]);
  const uint32_t t = threadIdx.x;
  for (uint32_t w = blockIdx.x; w < n; w += gridDim.x) {
    const uint2 item = D.dlist[6][w];
    const uint32_t chunk = (uint32_t)__builtin_amdgcn_readfirstlane((int)item.y);
    if (chunk != 0u) continue;   // one workgroup per tile
    const uint2 tile = D.tileAll[(uint32_t)__builtin_amdgcn_readfirstlane((int)item.x)];
    const uint32_t c0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)tile.x);
    const uint32_t yy = (uint32_t)__builtin_amdgcn_readfirstlane((int)tile.y);
    if ((yy & 0xffu) != 1u) continue;
    const SceneP& P = D.sp[yy >> 8];
    const uint32_t lo = (uint32_t)((int32_t)D.cellStart[c0] + P.posShift);
    const uint32_t hi = (uint32_t)((int32_t)D.cellStart[c0 + 1] + P.posShift);
    if (hi - lo <= TILE_CHUNK) continue;
    __syncthreads();   // the previous tile's bins are no longer read
    if (t < 64u) {
      bins[t] = 0u;
      blive[t] = 0u;
      bx0[t] = by0[t] = I32_MAX;
      bx1[t] = by1[t] = I32_MIN;
    }
    __syncthreads();
    for (uint32_t i = lo + t; i < hi; i += 256u) {
      const int2 pr = D.pred[i];
      const uint32_t b = morton_bin(P, c0, pr);
      atomicAdd(&bins[b], 1u);
      if (!(D.info[i] & INF_DEAD)) {
        atomicAdd(&blive[b], 1u);
        atomicMin(&bx0[b], pr.x);
        atomicMax(&bx1[b], pr.x);
        atomicMin(&by0[b], pr.y);
        atomicMax(&by1[b], pr.y);
      }
    }
    __syncthreads();
    if (t < 64u) {   // exclusive scan of the 64 bin counts, one wave
      const uint32_t v = bins[t];
      uint32_t incl = v;
#pragma unroll
      for (uint32_t off = 1; off < 64u; off <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)incl, off);
        if (t >= off) incl += u;
      }
      bins[t] = incl - v;
      bstart[t] = incl - v;
    }
    if (t == 0u) slotS = atomicAdd(&D.listCount[8], 1u);
    __syncthreads();
    for (uint32_t i = lo + t; i < hi; i += 256u)
      D.hperm[lo + atomicAdd(&bins[morton_bin(P, c0, D.pred[i])], 1u)] = i;
    __syncthreads();   // bins[] now holds each bin's end
    const uint32_t slot = slotS;
    if (slot < D.binCap) {
      if (t < 64u) {
        D.binTab[(size_t)slot * 64u + t] = make_uint4(bstart[t], bins[t], blive[t], 0u);
        D.binBox[(size_t)slot * 64u + t] = make_int4(bx0[t], by0[t], bx1[t], by1[t]);
      }
      if (t == 0u) D.cellBin[c0] = slot;
    }
  }
}

template <bool STRIP>
__device__ __forceinline__ void nbr_item(const Dev& D, uint2 item, uint16_t* lst, uint4* stg, uint32_t lane,
                                         int32_t W0, bool heavy) {
  const uint32_t tix = __builtin_amdgcn_readfirstlane(item.x);
  const uint32_t chunk = __builtin_amdgcn_readfirstlane(item.y);
  const uint2 tile = D.tileAll[tix];
  const uint32_t c0 = __builtin_amdgcn_readfirstlane(tile.x);
  const uint32_t yy = __builtin_amdgcn_readfirstlane(tile.y);
  const uint32_t ncell = yy & 0xffu, s = yy >> 8;
  const bool bnd = D.sp[s].bounded != 0u;
  const int32_t sh = D.sp[s].posShift;
  const uint32_t tlo = (uint32_t)((int32_t)D.cellStart[c0] + sh);
  const uint32_t thi = (uint32_t)((int32_t)D.cellStart[c0 + ncell] + sh);
  const uint32_t plo = tlo + chunk * TILE_CHUNK;
  const uint32_t phi = (uint32_t)mini((int32_t)thi, (int32_t)(plo + TILE_CHUNK));
  const bool one = (uint32_t)__builtin_amdgcn_readfirstlane((int)(phi - plo)) <= 64u;
  // a packed one-cell tile of the heavy list takes its chunks in the spatial
  // order k_heavy_order built, so each wave's particles lie close together
  const uint32_t* perm = (heavy && ncell == 1u && thi - tlo > TILE_CHUNK) ? D.hperm : nullptr;
  Dbg g;
  if (bnd) {
    // one-cell tiles of cells no wider than 4 wu: every particle's own
    // stencil is the whole candidate union (no window) and |d| < 2^19
    if (ncell == 1u && D.sp[s].cellShift <= 18u) {
      if (!(one && nbr_chunk_staged<true, false, STRIP>(D, lst, stg, c0, ncell, s, plo, phi, lane, W0, g, perm)))
        nbr_chunk<true, false, NBR_PPL, STRIP>(D, stg, c0, ncell, s, plo, phi, lane, W0, g, perm);
    } else {
      if (!(one && nbr_chunk_staged<true, true, STRIP>(D, lst, stg, c0, ncell, s, plo, phi, lane, W0, g, perm)))
        nbr_chunk<true, true, NBR_PPL, STRIP>(D, stg, c0, ncell, s, plo, phi, lane, W0, g, perm);
    }
  } else {
    nbr_chunk<false, true, NBR_PPL, STRIP>(D, stg, c0, ncell, s, plo, phi, lane, W0, g, perm);
  }
  flush_dbg(D, s, g, 0);
}

// STRIP: the build without debug counters (a full list ends a particle's
// sweep, with the per-particle tails); the debug build's kernel carries
// none of that code, so its register budget is its own.
template <bool STRIP>
__global__ __launch_bounds__(256, 5) void k_nbr_tile(Dev D) {
  __shared__ uint16_t lists[4][(MAXNBR + 1) * 64];   // a spare row per wave: see nbr_chunk_staged
  __shared__ uint4 stage[4][64];                     // a block's candidates per wave: see nbr_chunk_staged
  const uint32_t lane = threadIdx.x & 63u;
  uint16_t* lst = lists[threadIdx.x >> 6];
  uint4* stg = stage[threadIdx.x >> 6];
  const int32_t W0 = D.lut4[0].x;
  // heavy tiles first, each chunk taken by the next free wave anywhere
  const uint32_t nheavy = (uint32_t)__builtin_amdgcn_readfirstlane((int)D.listCount[6]);
  while (nheavy) {
    uint32_t w = 0;
    if (lane == 0u) w = atomicAdd(&D.listCount[7], 1u);
    w = (uint32_t)__builtin_amdgcn_readfirstlane((int)w);
    if (w >= nheavy) break;
    nbr_item<STRIP>(D, D.dlist[6][w], lst, stg, lane, W0, true);
  }
  const uint32_t nwork = (uint32_t)__builtin_amdgcn_readfirstlane((int)D.listCount[4]);
  const ListRange R = xcd_range(nwork, 4u, threadIdx.x >> 6);
  for (uint32_t w = R.first; w < R.hi; w += R.stride) nbr_item<STRIP>(D, D.dlist[4][w], lst, stg, lane, W0, false);
}

// ================================================================ solveA (k>0)

template <bool BND>
__device__ __forceinline__ void solveA_at(const Dev& D, uint32_t p, int2 pi, uint32_t cnt, const int2* lutWG,
                                          Dbg& g) {
  const uint32_t* nb = D.nbr + p;
  SolveAAcc acc{D.lut4[0].x, 0, 0, 0};
  // the list words of each group are loaded while the group before it is
  // gathered and summed (list slots below MAXNBR are always inside the array)
  uint32_t qn[NBR_UNROLL];
#pragma unroll
  for (uint32_t u = 0; u < NBR_UNROLL; ++u) qn[u] = nb[(size_t)u * D.NS];
  for (uint32_t k = 0; k < cnt; k += NBR_UNROLL) {
    uint32_t q[NBR_UNROLL];
    int2 d[NBR_UNROLL];
    int32_t idx[NBR_UNROLL];
    int4 L[NBR_UNROLL];
#pragma unroll
    for (uint32_t u = 0; u < NBR_UNROLL; ++u) q[u] = k + u < cnt ? qn[u] : p;
    if (k + NBR_UNROLL < MAXNBR) {
#pragma unroll
      for (uint32_t u = 0; u < NBR_UNROLL; ++u) qn[u] = nb[(size_t)(k + NBR_UNROLL + u) * D.NS];
    }
#pragma unroll
    for (uint32_t u = 0; u < NBR_UNROLL; ++u) {
      d[u] = pdelta<BND>(pi, D.pred[q[u] & ~NBR_BND], g);
      idx[u] = (k + u < cnt) ? lut_idx<BND>(d[u], g) : -1;
      const int2 wg = lutWG[maxi(idx[u], 0)];
      L[u] = make_int4(wg.x, wg.y, 0, 0);   // solveA uses W and G only
    }
#pragma unroll
    for (uint32_t u = 0; u < NBR_UNROLL; ++u)
      if (idx[u] >= 0) solveA_term<BND>(acc, d[u], L[u], g);
  }
  solveA_finish(D, p, pi, acc, MAT(D, (uint32_t)D.ptY[p].x)[M_EPS], g);
}

// True on every lane when the wave's active lanes all belong to bounded scenes.
__device__ __forceinline__ bool wave_bounded(const Dev& D, uint32_t s) {
  return __all(s == INACTIVE || D.sp[s].bounded != 0u);
}

__global__ __launch_bounds__(256) void k_solveA(Dev D) {
  // the (W, G) kernel tables in LDS: every pair term reads one entry
  __shared__ int2 lutWG[TN];
  for (uint32_t i = threadIdx.x; i < TN; i += 256u) {
    const int4 L = D.lut4[i];
    lutWG[i] = make_int2(L.x, L.y);
  }
  __syncthreads();
  const uint32_t p = slot_block(D) * 256u + threadIdx.x;
  // the particle's own words are loaded with its scene index (one round trip)
  uint32_t s = INACTIVE, inf = 0, nn = 0;
  int2 pi = make_int2(0, 0);
  if (p < D.N) {
    s = D.sidOf[p];
    inf = D.info[p];
    nn = D.nbrN[p];
    pi = D.pred[p];
  }
  Dbg g;
  const bool bnd = wave_bounded(D, s);
  if (s != INACTIVE && (inf & INF_FLUID)) {
    if (bnd) solveA_at<true>(D, p, pi, nn & 0x1fffffffu, lutWG, g);
    else solveA_at<false>(D, p, pi, nn & 0x1fffffffu, lutWG, g);
  }
  flush_dbg(D, s, g, 0);
}

// A colour-list item {first sorted slot, count | scene << COL_SCENE_SHIFT},
// read uniformly by the workgroup.
struct ColItem {
  uint32_t lo, m, s;
};
__device__ __forceinline__ ColItem col_item(uint2 item) {
  ColItem c;
  c.lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)item.x);
  const uint32_t y = (uint32_t)__builtin_amdgcn_readfirstlane((int)item.y);
  c.m = y & COL_COUNT;
  c.s = y >> COL_SCENE_SHIFT;
  return c;
}


// ================================================================ colour pass
// solveB_c then applyDp_c (src/sim.wgsl:1337-1474) for one colour, as two
// launches over the colour's cell chunks, like the reference's two
// dispatches: every dp of the colour is formed before any pred moves.
// k_colour gives each chunk one workgroup and splits each particle's
// neighbour list over L lanes (entries k = sub, sub + L, ...; L = 256 /
// particles, a power of two up to 16) whose wrapping partial sums are
// combined with xor shuffles, so the 48-deep gather chains run L-fold
// shorter; the (G, S) kernel tables sit in LDS.  k_apply moves the chunk's
// particles and, on the last iteration, finalizes them
// (src/sim.wgsl:2245-2265).  (Moving a one-chunk cell's particles inside
// k_colour after a workgroup barrier is race-free, since a cell's only
// same-colour neighbour is itself, but measured slower on the 1M dam break:
// the barrier idles the workgroup, 3.02 against 2.96 ms per substep.)
// A lane takes entries k = sub + u * L, u < SB_E, of its particle's list
// (L >= 8 lanes per particle, so SB_E = 6 covers MAXNBR).  All list words
// are loaded first, then all neighbour preds and lambdas, then the terms
// are formed: one gather round trip per item instead of one per pair of
// entries.  Entries past the count are skipped before any test.
constexpr uint32_t SB_E = MAXNBR / 8;
static_assert(COLOUR_THREADS / CELL_CHUNK >= 8u && SB_E * 8u >= MAXNBR, "k_colour gives every particle at least 8 lanes");
static_assert(CELL_CHUNK <= COL_COUNT, "an item fits the count field");
static_assert(2ll * LAM_MAX <= (1ll << 23), "a zero-S solveB coefficient li + lj stays within 2^23");
// S0: the world's S table is all zero (the frozen image's is), so the LDS
// table holds G alone (12.8 KB instead of 25.6 KB: more workgroups per CU
// to hide the gathers).  The list words are loaded without waiting for the
// count (every k < MAXNBR is inside the list array) and replaced by the
// particle itself past it, so the count, the lists and the particle's own
// data arrive in one round trip.
template <bool BND, bool S0>
__device__ __forceinline__ int2 solveB_part(const Dev& D, uint32_t p, uint32_t sub, uint32_t L,
                                            const int32_t* lutT, Dbg& g) {
  const uint32_t* nb = D.nbr + p;
  uint32_t e[SB_E];
#pragma unroll
  for (uint32_t u = 0; u < SB_E; ++u) {
    const uint32_t k = min(sub + u * L, MAXNBR - 1u);
    e[u] = nb[(size_t)k * D.NS];
  }
  const int4 own = D.pl4[p];
  const int2 pi = make_int2(own.x, own.y);
  const int32_t li = own.z;
  const uint32_t cnt = D.nbrN[p] & 0x1fffffffu;
  int2 pj[SB_E];
  int32_t lj[SB_E];
#pragma unroll
  for (uint32_t u = 0; u < SB_E; ++u) {
    if (sub + u * L >= cnt) e[u] = p;
    const int4 r = D.pl4[e[u] & ~NBR_BND];   // pred and lambda in one gather
    pj[u] = make_int2(r.x, r.y);
    lj[u] = r.z;
  }
  int32_t dpx = 0, dpy = 0;
#pragma unroll
  for (uint32_t u = 0; u < SB_E; ++u) {
    if (sub + u * L >= cnt) continue;
    const int2 d = pdelta<BND>(pi, pj[u], g);
    const int32_t ix = lut_idx<BND>(d, g);
    if (ix < 0) continue;
    const int2 T = S0 ? make_int2(lutT[ix], 0) : reinterpret_cast<const int2*>(lutT)[ix];
    if constexpr (BND && S0) {
      // with a zero S table: |co| <= 2 LAM_MAX, so the sum cannot saturate,
      // and |co g| / 2^12 < 2^23 * 2^18.33 / 2^12 < 2^31, so neither can the
      // term (no event is lost: fp_add_sat and q_sat_clamp would count none)
      const int32_t co = li + ((e[u] & NBR_BND) ? li : lj[u]);
      const double gs = (double)T.x * 0x1p-26, cf = (double)co * 0x1p-12;
      const double gx = q_rne(gs * (double)d.x), gy = q_rne(gs * (double)d.y);
      dpx = wadd(dpx, (int32_t)__builtin_fmin(__builtin_fmax(q_rne(cf * gx), (double)-DP_TERM), (double)DP_TERM));
      dpy = wadd(dpy, (int32_t)__builtin_fmin(__builtin_fmax(q_rne(cf * gy), (double)-DP_TERM), (double)DP_TERM));
      continue;
    }
    const int32_t co = fp_add_sat(fp_add_sat(li, (e[u] & NBR_BND) ? li : lj[u], g), T.y, g);
    if constexpr (BND) {
      // binary64 forms (bounds as in solveA_term): exact gradients with
      // |g| < 2^18.33, then |co g| < 2^31 * 2^18.33 < 2^53 is exact too
      const double gs = (double)T.x * 0x1p-26, cf = (double)co * 0x1p-12;
      const double gx = q_rne(gs * (double)d.x), gy = q_rne(gs * (double)d.y);
      dpx = wadd(dpx, q_sat_clamp(q_rne(cf * gx), -DP_TERM, DP_TERM, g));
      dpy = wadd(dpy, q_sat_clamp(q_rne(cf * gy), -DP_TERM, DP_TERM, g));
    } else {
      dpx = wadd(dpx, clampi(fp_mul_s(co, grad_q16(T.x, d.x, g), 20 - ACC, g), -DP_TERM, DP_TERM));
      dpy = wadd(dpy, clampi(fp_mul_s(co, grad_q16(T.x, d.y, g), 20 - ACC, g), -DP_TERM, DP_TERM));
    }
  }
  return make_int2(dpx, dpy);
}

// applyDp for fluid particle p of scene P (src/sim.wgsl:1458-1474), then on
// the last iteration finalize (src/sim.wgsl:2245-2265).
__device__ __forceinline__ void apply_at(const Dev& D, const SceneP& P, uint32_t p, int2 dp, uint32_t last,
                                         Dbg& g) {
  const int4 b = D.ptY[p];
  const uint32_t fl = (uint32_t)b.y;
  int2 pr = D.pred[p];
  if ((fl & (FLAG_RIGID | FLAG_ANCHOR)) == 0u) {
    const int32_t* M = MAT(D, (uint32_t)b.x);
    const int32_t dm = M[M_DPMAX], dmN = fp_neg_sat(dm, g);
    const int32_t ddx = clampi(fp_shr_rne(dp.x, ACC), dmN, dm);
    const int32_t ddy = clampi(fp_shr_rne(dp.y, ACC), dmN, dm);
    pr = wall_clamp(P, make_int2(fp_add_sat(pr.x, ddx, g), fp_add_sat(pr.y, ddy, g)));
    D.pred[p] = pr;
    *reinterpret_cast<int2*>(&D.pl4[p]) = pr;   // the record's pred half
  }
  if (last) {
    const int4 a = D.psY[p];
    const int32_t vmaxN = fp_neg_sat(P.vmax, g);
    const int32_t vx = clampi(fp_sub_sat(pr.x, a.x, g), vmaxN, P.vmax);
    const int32_t vy = clampi(fp_sub_sat(pr.y, a.y, g), vmaxN, P.vmax);
    const int32_t rho = D.rho[p];
    const int32_t ri = clampi(fp_divshift(RHO0, maxi(rho, RHO_FLOOR_RINV), 16, g), 0, RINV_MAX);
    D.vtmp[p] = make_int2(vx, vy);
    D.rinv[p] = ri;
    D.vr4[p] = make_int4(vx, vy, rho, ri);
  }
}

template <bool S0>
__global__ __launch_bounds__(COLOUR_THREADS) void k_colour(Dev D, uint32_t colour) {
  __shared__ int32_t lutT[S0 ? TN : 2 * TN];   // G, or (G, S) pairs
  for (uint32_t i = threadIdx.x; i < TN; i += COLOUR_THREADS) {
    const int4 L = D.lut4[i];
    if (S0) lutT[i] = L.y;
    else reinterpret_cast<int2*>(lutT)[i] = make_int2(L.y, L.w);
  }
  __syncthreads();
  const uint32_t n = (uint32_t)__builtin_amdgcn_readfirstlane((int)D.listCount[colour]);
  const uint2* list = D.dlist[colour];
  const ListRange R = xcd_range(n, 1u, 0u);
  for (uint32_t w = R.first; w < R.hi; w += R.stride) {
    const ColItem it = col_item(list[w]);
    const bool bnd = D.sp[it.s].bounded != 0u;
    Dbg g;
    uint32_t L = 16;
    while (L > 1u && L * it.m > COLOUR_THREADS) L >>= 1;
    const uint32_t pl = threadIdx.x / L, sub = threadIdx.x & (L - 1u);
    const uint32_t p = it.lo + pl;
    const bool act = pl < it.m && (D.info[p] & INF_FLUID);
    int2 dp = make_int2(0, 0);
    if (act) dp = bnd ? solveB_part<true, S0>(D, p, sub, L, lutT, g) : solveB_part<false, S0>(D, p, sub, L, lutT, g);
    for (uint32_t off = 1; off < L; off <<= 1) {
      dp.x = wadd(dp.x, __shfl_xor(dp.x, (int)off));
      dp.y = wadd(dp.y, __shfl_xor(dp.y, (int)off));
    }
    if (act && sub == 0u) D.dp[p] = dp;
    flush_dbg(D, it.s, g, 0);
  }
}

// applyDp of the colour's chunks, one wave per chunk.
__global__ __launch_bounds__(256) void k_apply(Dev D, uint32_t colour, uint32_t last) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t n = (uint32_t)__builtin_amdgcn_readfirstlane((int)D.listCount[colour]);
  const uint2* list = D.dlist[colour];
  const ListRange R = xcd_range(n, 4u, threadIdx.x >> 6);
  for (uint32_t w = R.first; w < R.hi; w += R.stride) {
    const ColItem it = col_item(list[w]);
    Dbg g;
    const uint32_t p = it.lo + lane;
    if (lane < it.m && (D.info[p] & INF_FLUID)) apply_at(D, D.sp[it.s], p, D.dp[p], last, g);
    flush_dbg(D, it.s, g, 0);
  }
}

// ========================================================= xsph+normals+conduct
// src/sim.wgsl:2300-2384 and :2528-2582 share one gather over fluid
// neighbours; each keeps its own accumulators.  Writes the final position and
// velocity of the substep (boundary slots: position kept, velocity zeroed).
template <bool BND>
__device__ __forceinline__ void xnc_fluid(const Dev& D, uint32_t p, const int2* lutWG, Dbg& g) {
  int4 b = D.ptY[p];
  uint32_t mi = (uint32_t)b.x;
  const int32_t* Mi = MAT(D, mi);
  uint32_t fl = (uint32_t)b.y;
  bool doX = (fl & (FLAG_RIGID | FLAG_ANCHOR)) == 0u;
  int32_t ci = Mi[M_XSPH], ki = Mi[M_COND], Ti = b.z, corri = Mi[M_CORRODE];
  int2 pi = D.pred[p], vi = D.vtmp[p];
  int32_t rhoi = D.rho[p];
  uint32_t cnt = D.nbrN[p] & 0x1fffffffu;
  const uint32_t* nb = D.nbr + p;
  int32_t ax = 0, ay = 0, sx = 0, sy = 0, q = 0, c = 0;
  // the list words of each group are loaded while the group before it is
  // gathered and summed (list slots below MAXNBR are always inside the array)
  uint32_t en[XNC_UNROLL];
#pragma unroll
  for (uint32_t u = 0; u < XNC_UNROLL; ++u) en[u] = nb[(size_t)u * D.NS];
  for (uint32_t k = 0; k < cnt; k += XNC_UNROLL) {
    // padding entries read as boundary entries, which this loop skips
    uint32_t e[XNC_UNROLL];
    int2 d[XNC_UNROLL], vjs[XNC_UNROLL];
    int32_t idx[XNC_UNROLL], rhoj[XNC_UNROLL], rjs[XNC_UNROLL];
    int4 Ls[XNC_UNROLL], bjs[XNC_UNROLL];
#pragma unroll
    for (uint32_t u = 0; u < XNC_UNROLL; ++u) e[u] = k + u < cnt ? en[u] : (p | NBR_BND);
    if (k + XNC_UNROLL < MAXNBR) {
#pragma unroll
      for (uint32_t u = 0; u < XNC_UNROLL; ++u) en[u] = nb[(size_t)(k + XNC_UNROLL + u) * D.NS];
    }
#pragma unroll
    for (uint32_t u = 0; u < XNC_UNROLL; ++u) {
      // a skipped (boundary or padding) entry is measured against the
      // particle itself: a zero offset, no saturation event
      const bool skip = (e[u] & NBR_BND) != 0u;
      const uint32_t qj = e[u] & ~NBR_BND;
      d[u] = pdelta<BND>(pi, skip ? pi : D.pred[qj], g);
      idx[u] = skip ? -1 : lut_idx<BND>(d[u], g);
      const int2 wg = lutWG[maxi(idx[u], 0)];
      Ls[u] = make_int4(wg.x, wg.y, 0, 0);   // xsph, normals and conduct use W and G only
      bjs[u] = D.ptY[qj];
      const int4 vr = D.vr4[qj];   // velocity, rho and rinv in one gather
      vjs[u] = make_int2(vr.x, vr.y);
      rhoj[u] = vr.z;
      rjs[u] = vr.w;
    }
#pragma unroll
    for (uint32_t u = 0; u < XNC_UNROLL; ++u) {
      if (idx[u] < 0) continue;
      const int4 L = Ls[u], bj = bjs[u];
      const int32_t* Mj = MAT(D, (uint32_t)bj.x);
      if constexpr (BND) {
        // binary64 forms; the bounds that make each product exact hold in
        // bounded scenes of worlds whose tables and materials pass
        // pair_bounds (checked when the world is created)
        if (doX) {
          // |W| < 2^18, 0 <= K <= 2^18: w exact, |w| < 2^20; |dv| <= 2^31:
          // |dv w| < 2^51; the saturated t times |cij| < 2^22: < 2^53
          const int32_t K = sym_k(rhoi, rhoj[u], g);
          const double wf = q_rne((double)L.x * ((double)K * 0x1p-16)) * 0x1p-12;
          const double cf = (double)mat_mix(ci, Mj[M_XSPH]) * 0x1p-16;
          const int2 vj = vjs[u];
          const double tx = (double)q_sat(q_rne((double)fp_sub_sat(vj.x, vi.x, g) * wf), g);
          const double ty = (double)q_sat(q_rne((double)fp_sub_sat(vj.y, vi.y, g) * wf), g);
          ax = wadd(ax, q_sat_clamp(q_rne(tx * cf), -XS_TERM, XS_TERM, g));
          ay = wadd(ay, q_sat_clamp(q_rne(ty * cf), -XS_TERM, XS_TERM, g));
        }
        // exact gradients (|g| < 2^18.33) times rinv <= 2^18: < 2^36.33
        const double gs = (double)L.y * 0x1p-26, rf = (double)rjs[u] * 0x1p-16;
        sx = wadd(sx, (int32_t)q_rne(q_rne(gs * (double)d[u].x) * rf));
        sy = wadd(sy, (int32_t)q_rne(q_rne(gs * (double)d[u].y) * rf));
        // |W| < 2^18, |cond| < 2^26: |wk| < 2^28; |dT| <= 2^24: < 2^52
        const double wk = q_rne((double)L.x * ((double)mat_mix(ki, Mj[M_COND]) * 0x1p-16));
        const int32_t dT = clampi(fp_sub_sat(bj.z, Ti, g), -DT_CLAMP, DT_CLAMP);
        q = wadd(q, q_sat_clamp(q_rne((double)dT * wk * 0x1p-12), -HT_TERM, HT_TERM, g));
        if ((uint32_t)bj.x != mi) {
          const int32_t attack = Mj[M_CORRODE];
          const int32_t spend = (Mj[M_CORRPT] != CORR_NEVER) ? corri : 0;
          const int32_t str = wadd(attack, spend);
          // |W| < 2^18, |str| <= 2^31: < 2^49
          if (str != 0) c = wadd(c, q_sat_clamp(q_rne((double)L.x * ((double)str * 0x1p-12)), 0, CORR_TERM, g));
        }
        continue;
      }
      if (doX) {
        int32_t K = sym_k(rhoi, rhoj[u], g);
        int32_t w = fp_mul_s(L.x, K, 16, g);
        int32_t cij = mat_mix(ci, Mj[M_XSPH]);
        int2 vj = vjs[u];
        ax = wadd(ax, clampi(fp_mul_s(fp_mul_s(fp_sub_sat(vj.x, vi.x, g), w, 20 - ACC, g), cij, 16, g), -XS_TERM, XS_TERM));
        ay = wadd(ay, clampi(fp_mul_s(fp_mul_s(fp_sub_sat(vj.y, vi.y, g), w, 20 - ACC, g), cij, 16, g), -XS_TERM, XS_TERM));
      }
      int32_t rj = rjs[u];
      sx = wadd(sx, fp_mul_s(grad_q16(L.y, d[u].x, g), rj, 16, g));
      sy = wadd(sy, fp_mul_s(grad_q16(L.y, d[u].y, g), rj, 16, g));
      int32_t cij = mat_mix(ki, Mj[M_COND]);
      int32_t wk = fp_mul_s(L.x, cij, 16, g);
      int32_t dT = clampi(fp_sub_sat(bj.z, Ti, g), -DT_CLAMP, DT_CLAMP);
      q = wadd(q, clampi(fp_mul_s(dT, wk, 20 - ACC, g), -HT_TERM, HT_TERM));
      if ((uint32_t)bj.x != mi) {
        int32_t attack = Mj[M_CORRODE];
        int32_t spend = (Mj[M_CORRPT] != CORR_NEVER) ? corri : 0;
        int32_t str = wadd(attack, spend);
        if (str != 0) c = wadd(c, clampi(fp_mul_s(L.x, str, 20 - ACC, g), 0, CORR_TERM));
      }
    }
  }
  int32_t vx = vi.x, vy = vi.y;
  if (doX) {
    vx = fp_add_sat(vx, fp_shr_rne(ax, ACC), g);
    vy = fp_add_sat(vy, fp_shr_rne(ay, ACC), g);
  }
  D.ps[p] = make_int4(pi.x, pi.y, vx, vy);
  const int2 nrm = make_int2(fp_mul_s(sx, H, 16, g), fp_mul_s(sy, H, 16, g));
  D.nrm[p] = nrm;
  D.nr4[p] = make_int4(nrm.x, nrm.y, rhoi, (int32_t)mi);   // surface tension's gather record
  D.heat[p] = q;
  D.corr[p] = c;
}

__global__ __launch_bounds__(256) void k_xnc(Dev D) {
  // the (W, G) kernel tables in LDS: every pair term reads one entry
  __shared__ int2 lutWG[TN];
  for (uint32_t i = threadIdx.x; i < TN; i += 256u) {
    const int4 L = D.lut4[i];
    lutWG[i] = make_int2(L.x, L.y);
  }
  __syncthreads();
  uint32_t p = slot_block(D) * 256u + threadIdx.x;
  uint32_t s = p < D.N ? D.sidOf[p] : INACTIVE;
  Dbg g;
  const bool bnd = wave_bounded(D, s);
  if (s != INACTIVE) {
    uint32_t inf = D.info[p];
    if (!(inf & INF_FLUID)) {
      int4 a = D.psY[p];
      D.ps[p] = make_int4(a.x, a.y, 0, 0);
    } else if (bnd) {
      xnc_fluid<true>(D, p, lutWG, g);
    } else {
      xnc_fluid<false>(D, p, lutWG, g);
    }
  }
  flush_dbg(D, s, g, 0);
}

// surfaceTension's pair sum (src/sim.wgsl:2409-2494) for fluid particle p of
// material row M: cohesion and curvature against fluid neighbours, adhesion
// against boundary ones.
template <bool BND>
__device__ __forceinline__ int2 st_sum(const Dev& D, uint32_t p, const int32_t* M, const int2* lutAC, Dbg& g) {
  int2 pi = D.pred[p], ni = D.nrm[p];
  int32_t rhoi = D.rho[p];
  uint32_t cnt = D.nbrN[p] & 0x1fffffffu;
  const uint32_t* nb = D.nbr + p;
  int32_t ax = 0, ay = 0;
  // list words one group ahead, as in xnc_fluid
  uint32_t en[XNC_UNROLL];
#pragma unroll
  for (uint32_t u = 0; u < XNC_UNROLL; ++u) en[u] = nb[(size_t)u * D.NS];
  for (uint32_t k = 0; k < cnt; k += XNC_UNROLL) {
    uint32_t e[XNC_UNROLL];
    int2 d[XNC_UNROLL], njs[XNC_UNROLL];
    int32_t idx[XNC_UNROLL], rhoj[XNC_UNROLL], cts[XNC_UNROLL], ats[XNC_UNROLL], matj[XNC_UNROLL];
#pragma unroll
    for (uint32_t u = 0; u < XNC_UNROLL; ++u) e[u] = k + u < cnt ? en[u] : p;
    if (k + XNC_UNROLL < MAXNBR) {
#pragma unroll
      for (uint32_t u = 0; u < XNC_UNROLL; ++u) en[u] = nb[(size_t)(k + XNC_UNROLL + u) * D.NS];
    }
#pragma unroll
    for (uint32_t u = 0; u < XNC_UNROLL; ++u) {
      uint32_t qj = e[u] & ~NBR_BND;
      d[u] = pdelta<BND>(pi, D.pred[qj], g);
      idx[u] = (k + u < cnt) ? lut_idx<BND>(d[u], g) : -1;
      uint32_t ii = (uint32_t)maxi(idx[u], 0);
      const int2 ac = lutAC[ii];
      ats[u] = ac.x;
      cts[u] = ac.y;
      const int4 nr = D.nr4[qj];   // normal, rho and material in one gather (fluid neighbours)
      rhoj[u] = nr.z;
      matj[u] = nr.w;
      njs[u] = make_int2(nr.x, nr.y);
    }
#pragma unroll
    for (uint32_t u = 0; u < XNC_UNROLL; ++u) {
    if (idx[u] < 0) continue;
    const int2 d1 = d[u];
    if constexpr (BND) {
      // binary64 forms (bounds as in xnc_fluid): |d| < 2^17.33 per axis
      if (e[u] & NBR_BND) {
        // |A d| < 2^48.33, exact with |a| < 2^26.33; times |adh| < 2^22: < 2^48.33
        const double as = (double)ats[u] * 0x1p-22, hf = (double)M[M_ADH] * 0x1p-8;
        ax = wsub(ax, q_sat_clamp(q_rne(q_rne(as * (double)d1.x) * hf), -ST_TERM, ST_TERM, g));
        ay = wsub(ay, q_sat_clamp(q_rne(q_rne(as * (double)d1.y) * hf), -ST_TERM, ST_TERM, g));
        continue;
      }
      const int32_t K = sym_k(rhoi, rhoj[u], g);
      const int32_t* Mj = MAT(D, (uint32_t)matj[u]);
      // |C d| < 2^48.33 exact with |c| < 2^26.33; times |gCoh| < 2^22: < 2^48.33
      const double cs = (double)cts[u] * 0x1p-22;
      const double hf = (double)mat_mix(M[M_GCOH], Mj[M_GCOH]) * 0x1p-16;
      int32_t fx = fp_neg_sat(q_sat(q_rne(q_rne(cs * (double)d1.x) * hf), g), g);
      int32_t fy = fp_neg_sat(q_sat(q_rne(q_rne(cs * (double)d1.y) * hf), g), g);
      // |dn| <= 2^31 times |gCur| < 2^22: < 2^53
      const int2 nj = njs[u];
      const double uf = (double)mat_mix(M[M_GCUR], Mj[M_GCUR]) * 0x1p-16;
      const int32_t cx = q_sat(q_rne((double)fp_sub_sat(ni.x, nj.x, g) * uf), g);
      const int32_t cy = q_sat(q_rne((double)fp_sub_sat(ni.y, nj.y, g) * uf), g);
      fx = fp_add_sat(fx, fp_neg_sat(cx, g), g);
      fy = fp_add_sat(fy, fp_neg_sat(cy, g), g);
      // |f| <= 2^31 times 0 <= K <= 2^18: < 2^49
      const double kf = (double)K * 0x1p-8;
      ax = wadd(ax, q_sat_clamp(q_rne((double)fx * kf), -ST_TERM, ST_TERM, g));
      ay = wadd(ay, q_sat_clamp(q_rne((double)fy * kf), -ST_TERM, ST_TERM, g));
      continue;
    }
    if (e[u] & NBR_BND) {
      int32_t at = ats[u];
      ax = wsub(ax, clampi(fp_mul_s(fp_mul_s(at, d1.x, 22, g), M[M_ADH], 16 - ACC, g), -ST_TERM, ST_TERM));
      ay = wsub(ay, clampi(fp_mul_s(fp_mul_s(at, d1.y, 22, g), M[M_ADH], 16 - ACC, g), -ST_TERM, ST_TERM));
      continue;
    }
    int32_t K = sym_k(rhoi, rhoj[u], g);
    int32_t ct = cts[u];
    const int32_t* Mj = MAT(D, (uint32_t)matj[u]);
    int32_t gCoh = mat_mix(M[M_GCOH], Mj[M_GCOH]);
    int32_t gCur = mat_mix(M[M_GCUR], Mj[M_GCUR]);
    int32_t fx = fp_neg_sat(fp_mul_s(fp_mul_s(ct, d1.x, 22, g), gCoh, 16, g), g);
    int32_t fy = fp_neg_sat(fp_mul_s(fp_mul_s(ct, d1.y, 22, g), gCoh, 16, g), g);
    int2 nj = njs[u];
    int32_t cx = fp_mul_s(fp_sub_sat(ni.x, nj.x, g), gCur, 16, g);
    int32_t cy = fp_mul_s(fp_sub_sat(ni.y, nj.y, g), gCur, 16, g);
    fx = fp_add_sat(fx, fp_neg_sat(cx, g), g);
    fy = fp_add_sat(fy, fp_neg_sat(cy, g), g);
    ax = wadd(ax, clampi(fp_mul_s(fx, K, 16 - ACC, g), -ST_TERM, ST_TERM));
    ay = wadd(ay, clampi(fp_mul_s(fy, K, 16 - ACC, g), -ST_TERM, ST_TERM));
    }
  }
  return make_int2(ax, ay);
}

// ==================================================================== hash
// One slot's terms of the state hash (src/hash/statehash.wgsl:83-159): four
// salted MurmurHash3 lanes over the 8 state words; a fluid slot with bonds
// folds in its bond count and the two order-free sums over its row (bn = 0
// leaves the lanes as they are).  hv[0..3] are the H_set terms, hv[4..7] the
// H_slot terms.
__device__ __forceinline__ void hash_terms(const uint32_t words[8], uint32_t myid, uint32_t bn, uint32_t com,
                                           uint32_t scom, uint32_t hv[8]) {
  const uint32_t SEED[4] = {0x9e3779b1u, 0x85ebca77u, 0xc2b2ae3du, 0x27d4eb2fu};
  const uint32_t SALT[4] = {0x00000000u, 0x51633e2du, 0x9c8f2a17u, 0xd0e5b943u};
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    uint32_t h = SEED[l];
#pragma unroll
    for (int w = 0; w < 8; ++w) h = mm_step(h, words[w] ^ SALT[l]);
    uint32_t hs = h, hl = h;
    if (bn != 0u) {
      hs = mm_step(h, bn) ^ com;
      hl = mm_step(h, bn) ^ scom;
    }
    hv[l] = fmix32(hs ^ 8u);
    hv[4 + l] = fmix32(mm_step(hl, myid) ^ 9u);
  }
}

// Block partial (8 hash words + truncation count): a block lying wholly in
// one scene stores its sums in hpart[block], which k_hash_finish folds in;
// a block straddling scenes adds per lane with atomics.  All are wrapping
// u32 sums, so the result does not depend on the split.  Every thread of
// the block must call it.
__device__ __forceinline__ void hash_reduce(const Dev& D, uint32_t blk, uint32_t s,
                                            const uint32_t hv[HPART_WORDS]) {
  __shared__ uint32_t wpart[4][HPART_WORDS];
  __shared__ uint32_t wscene[4];
  const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t s0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)s);
  const bool wuni = __all(s == s0);
#pragma unroll
  for (int l = 0; l < HPART_WORDS; ++l) {
    uint32_t v = wave_sum(hv[l]);
    if (lane == 0u) wpart[wv][l] = v;
  }
  if (lane == 0u) wscene[wv] = wuni ? s0 : MIXED;
  __syncthreads();
  const uint32_t bs = wscene[0];
  const bool buni = bs != MIXED && wscene[1] == bs && wscene[2] == bs && wscene[3] == bs;
  if (buni) {
    if (bs != INACTIVE && threadIdx.x < HPART_WORDS) {
      uint32_t t = threadIdx.x;
      D.hpart[(size_t)blk * HPART_WORDS + t] = wpart[0][t] + wpart[1][t] + wpart[2][t] + wpart[3][t];
    }
  } else if (s != INACTIVE) {
#pragma unroll
    for (int l = 0; l < 8; ++l) atomicAdd(&D.sd[s].acc[l], hv[l]);
    if (hv[8]) atomicAdd(&D.sd[s].dbg[6], hv[8]);
  }
}

// ==================================================== surfaceTension + thermal
// src/sim.wgsl:2409-2494 (stAcc for the next substep), :2612-2707 (thermal,
// elementwise), bondForm2's flag clear on formation substeps (:3086-3096;
// bond rows stay empty for bondless rows), then the particle's state hash
// (src/hash/statehash.wgsl:83-159) reduced per scene.  With gen = 1 (the
// material passes run) the flag clear and the hash are left to
// k_bond_form2 and k_hash_gen.
__global__ __launch_bounds__(256) void k_st_thermal_hash(Dev D, uint32_t gen) {
  // the (adhesion, cohesion) kernel tables in LDS: every pair term reads one entry
  __shared__ int2 lutAC[TN];
  for (uint32_t i = threadIdx.x; i < TN; i += 256u) lutAC[i] = make_int2(D.lutA[i], D.lut4[i].z);
  __syncthreads();
  const uint32_t blk = slot_block(D);
  uint32_t p = blk * 256u + threadIdx.x;
  uint32_t s = p < D.N ? D.sidOf[p] : INACTIVE;
  Dbg g;
  uint32_t words[8];
  uint32_t myid = 0;
  const bool bnd = wave_bounded(D, s);
  if (s != INACTIVE) {
    const SceneP& P = D.sp[s];
    const SceneD& SD = D.sd[s];
    bool formation = ((SD.substep % BOND_PERIOD) == 0u) || SD.bondDirty;
    uint32_t inf = D.info[p];
    int4 b = D.ptY[p];
    myid = D.idY[p];
    int4 a = D.ps[p];
    int2 stOut = D.stY[p];
    if (inf & INF_FLUID) {
      uint32_t mat0 = (uint32_t)b.x;
      const int32_t* M = MAT(D, mat0);
      // ---- surfaceTension
      const int2 axy = bnd ? st_sum<true>(D, p, M, lutAC, g) : st_sum<false>(D, p, M, lutAC, g);
      const int32_t ax = axy.x, ay = axy.y;
      stOut = make_int2(fp_shr_rne(ax, ACC), fp_shr_rne(ay, ACC));
      // ---- thermal
      int32_t T = fp_add_sat(b.z, fp_shr_rne(D.heat[p], ACC), g);
      T = fp_add_sat(T, fp_mul_s(fp_sub_sat(AMBIENT, T, g), M[M_COOL], 16, g), g);
      int32_t fuel = b.w;
      uint32_t fl0 = (uint32_t)b.y;
      bool burning = (fl0 & FLAG_BURNING) != 0u;
      if (!burning && fuel > 0 && T >= M[M_IGNITE]) burning = true;
      if (burning) {
        T = fp_add_sat(T, M[M_BURNHEAT], g);
        fuel = maxi(0, fp_sub_sat(fuel, M[M_BURNRATE], g));
      }
      T = clampi(T, TEMP_MIN, TEMP_MAX);
      T = fp_add_sat(T, fp_mul_s(fp_sub_sat(M[M_TEMPTARGET], T, g), M[M_TEMPDRIVE], 16, g), g);
      int32_t dose = fp_shr_rne(D.corr[p], ACC);
      uint32_t mat = mat0;
      bool froze = false;
      uint32_t splitTo = (uint32_t)M[M_SPLITTO], pct = (uint32_t)M[M_SPLITPCT];
      if (burning && fuel <= 0) mat = pick_product((uint32_t)M[M_BURNTO], splitTo, pct, myid);
      else if (dose >= M[M_CORRPT]) mat = pick_product((uint32_t)M[M_CORRTO], splitTo, pct, myid);
      else if (T >= M[M_BOILPT]) mat = pick_product((uint32_t)M[M_BOILTO], splitTo, pct, myid);
      else if (T >= M[M_MELTPT]) mat = pick_product((uint32_t)M[M_MELTTO], splitTo, pct, myid);
      else if (T <= M[M_CONDPT]) mat = pick_product((uint32_t)M[M_CONDTO], splitTo, pct, myid);
      else if (T <= M[M_FREEZEPT]) {
        mat = pick_product((uint32_t)M[M_FREEZETO], splitTo, pct, myid);
        froze = true;
      }
      const int32_t* Mn = MAT(D, mat);
      if (mat != mat0) {
        fuel = Mn[M_FUEL0];
        burning = false;
      }
      uint32_t fl = fl0 & ~(FLAG_BURNING | FLAG_DEAD | FLAG_SEAL);
      if (mat != mat0) fl &= ~(FLAG_ANCHOR | FLAG_CONGEAL | FLAG_BURST);
      if (froze && mat != mat0 && fuse_row(Mn, mat)) fl |= FLAG_CONGEAL;
      if (burning) fl |= FLAG_BURNING;
      if (Mn[M_PHANTOM] != 0) fl |= FLAG_DEAD;
      if (seal_class(Mn)) fl |= FLAG_SEAL;
      if (!gen && !supported_row(Mn)) atomicOr(&D.sd[s].fault, FAULT_MATERIAL);
      // ---- bondForm2's flag update on formation substeps (a world running
      // the material passes does it in k_bond_form2)
      if (formation && !gen) {
        fl &= ~FLAG_CONGEAL;
        if ((fl & FLAG_RIGID) == 0u && (fl & WELD_TAG_MASK) != P.weldHold) fl &= ~WELD_TAG_MASK;
      }
      b = make_int4((int32_t)mat, (int32_t)fl, T, fuel);
    }
    D.pt[p] = b;
    D.id[p] = myid;
    D.st[p] = stOut;
    words[0] = (uint32_t)a.x; words[1] = (uint32_t)a.y; words[2] = (uint32_t)a.z; words[3] = (uint32_t)a.w;
    words[4] = (uint32_t)b.x; words[5] = (uint32_t)b.y; words[6] = (uint32_t)b.z; words[7] = (uint32_t)b.w;
  }
  flush_dbg(D, s, g, 0);
  if (gen) return;   // hashed by k_hash_gen once the material passes have run
  // ---- state hash: 4 salted lanes over the 8 words, H_set and H_slot terms
  uint32_t hv[HPART_WORDS];
  if (s != INACTIVE) {
    hash_terms(words, myid, 0u, 0u, 0u, hv);
    hv[8] = D.drop[p];
  } else {
#pragma unroll
    for (int l = 0; l < HPART_WORDS; ++l) hv[l] = 0u;
  }
  hash_reduce(D, blk, s, hv);
}

#include "materials2d.hpp"

// ============================================================ stir events
// applyImpulse (src/engine.js:6077-6098) for scenes whose next scheduled event
// is due at their current substep.  k_impulse_mark flags hits in slot order;
// an exclusive scan over the flags ranks them; k_impulse_apply adds the
// impulse to the first `limit` hits of each scene.
__global__ __launch_bounds__(256) void k_impulse_mark(Dev D, uint32_t group) {
  uint32_t p = slot_block(D) * 256u + threadIdx.x;
  if (p >= D.N) return;
  uint32_t s = D.sidOf[p];
  if (s == INACTIVE) return;
  const SceneP& P = D.sp[s];
  const SceneD& SD = D.sd[s];
  uint32_t id = D.id[p];
  uint32_t slot = P.partBase + id;
  uint32_t hit = 0;
  uint32_t e = SD.evCursor + group;
  if (id < P.nFluid && e < SD.evBase + SD.evCount) {
    const Event& E = D.events[e];
    if (E.substep == SD.substep) {
      int4 a = D.ps[p];
      double dx = (double)a.x - (double)E.cx;
      double dy = (double)a.y - (double)E.cy;
      // IEEE double, no contraction (built with -ffp-contract=off), as JS
      double q = dx * dx;
      q = q + dy * dy;
      double r2 = (double)E.r * (double)E.r;
      hit = (q > r2) ? 0u : 1u;
    }
  }
  if (id < P.cap) D.hitFlag[slot] = hit;
}

__global__ __launch_bounds__(256) void k_impulse_apply(Dev D, uint32_t group) {
  uint32_t p = slot_block(D) * 256u + threadIdx.x;
  if (p >= D.N) return;
  uint32_t s = D.sidOf[p];
  if (s == INACTIVE) return;
  const SceneP& P = D.sp[s];
  const SceneD& SD = D.sd[s];
  uint32_t id = D.id[p];
  uint32_t e = SD.evCursor + group;
  if (id >= P.nFluid || e >= SD.evBase + SD.evCount) return;
  const Event& E = D.events[e];
  if (E.substep != SD.substep) return;
  uint32_t slot = P.partBase + id;
  if (!D.hitFlag[slot]) return;
  uint32_t rank = D.hitScan[slot] - D.hitScan[P.partBase];
  if (rank >= E.limit) return;
  int4 a = D.ps[p];
  int64_t vx = (int64_t)a.z + E.ix, vy = (int64_t)a.w + E.iy;
  vx = vx < -65536 ? -65536 : (vx > 65536 ? 65536 : vx);
  vy = vy < -65536 ? -65536 : (vy > 65536 ? 65536 : vy);
  a.z = (int32_t)vx;
  a.w = (int32_t)vy;
  D.ps[p] = a;
  atomicAdd(&D.sd[s].pushHits, 1u);   // a u32 count: order-free
}

// Advance each scene's event cursor past the group just applied.
__global__ void k_impulse_advance(Dev D, const uint8_t* active, uint32_t groups) {
  uint32_t s = blockIdx.x * 64u + threadIdx.x;
  if (s >= D.S || !active[s]) return;
  SceneD& SD = D.sd[s];
  uint32_t end = SD.evBase + SD.evCount;
  uint32_t k = 0;
  while (k < groups && SD.evCursor < end && D.events[SD.evCursor].substep == SD.substep) {
    SD.evCursor++;
    k++;
  }
  // events scheduled for a substep already passed are dropped in order
  while (SD.evCursor < end && D.events[SD.evCursor].substep < SD.substep) SD.evCursor++;
}

// Scatter the sorted state back to slot order for read-back.
__global__ __launch_bounds__(256) void k_to_slots(Dev D, uint32_t s, int32_t* st8, int32_t* dv16) {
  const SceneP& P = D.sp[s];
  uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= P.n) return;
  uint32_t p = P.partBase + t;
  uint32_t i = D.id[p];
  int4 a = D.ps[p], b = D.pt[p];
  int32_t* o = st8 + (size_t)i * 8;
  o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
  o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
  if (dv16) {
    int32_t* d = dv16 + (size_t)i * 16;
    int2 st = D.st[p];
    for (int k = 0; k < 16; ++k) d[k] = 0;
    d[0] = a.x; d[1] = a.y;     // pred == pos after finalize (boundary: always)
    d[8] = st.x; d[9] = st.y;
    if (i < P.nFluid && D.sd[s].substep > 0) {
      int2 vt = D.vtmp[p], nr = D.nrm[p];
      d[4] = vt.x; d[5] = vt.y;
      d[6] = nr.x; d[7] = nr.y;
      d[10] = D.rho[p]; d[11] = D.lam[p]; d[12] = D.rinv[p];
      d[13] = D.heat[p]; d[14] = D.corr[p];
    }
  }
}

inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

}  // namespace

// ========================================================================= host
struct sn_engine {
  sn_opts opts{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  Dev D{};
  // host mirrors
  std::vector<SceneP> sp;
  std::vector<SceneD> sdInit;
  std::vector<uint8_t> active;
  uint8_t* dActive = nullptr;
  std::vector<uint2> tiles;          // every neighbour tile of the active scenes
  uint32_t tileG = 1;
  uint32_t nbrBlocks = 0, colourBlocks = 0, applyBlocks = 0, canonBlocks = 0;  // persistent grid sizes
  uint32_t actBlocks = 1;            // entries of D.blockMap: the grid of the per-particle kernels
  uint32_t* dBlockMap = nullptr;
  uint32_t S = 0, N = 0, C = 0;
  uint32_t maxEventsPerSubstep = 0;
  Event* dEvents = nullptr;
  Event* dImpulse = nullptr;         // sn_apply_impulse's one-event schedule (allocated with the engine)
  uint32_t eventCap = 0;
  std::vector<void*> allocs;
  hipGraphExec_t graph = nullptr;
  uint32_t graphSubsteps = 0;
  bool loaded = false;
  bool slotted = false;  // sn_reserve_slots layout: partitions keep their bases
  uint32_t weldSeq = 1;  // the 12-bit spawn-event id of the last weld group (`_weldSeq`, src/engine.js:3801)
  bool gen = false;      // the material passes run (D.gen); never switched off again
  uint32_t graphG = 8;   // substeps per captured hipGraph (sn_autotune may change it)
  bool pairBounded = true;  // the tables and materials meet pair_bounds (bounded pair forms)
  bool sZero = true;        // the S kernel table is all zero (k_colour's LDS table holds G alone)
  // profiling: an event pair around every launch, tagged with its class
  std::vector<hipEvent_t> evPool;
  std::vector<int> evClass;
  uint32_t evUsed = 0;
  bool profiling = false;
  uint32_t profMask = ~0u;        // kernel classes timed while profiling (bit = ProfClass)
};

namespace {

template <class T>
T* dalloc(sn_engine* e, size_t count) {
  void* p = nullptr;
  HIPCHECK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
  e->allocs.push_back(p);
  return static_cast<T*>(p);
}

// chain start value (CHAIN_INIT, src/hash/statehash.js:28-31)
const uint32_t CHAIN_INIT[8] = {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u,
                                 0xa4093822u, 0x299f31d0u, 0x082efa98u, 0xec4e6c89u};

int fail(sn_engine* e, int code, const std::string& msg) {
  if (e) e->err = msg;
  return code;
}

void drop_graph(sn_engine* e) {
  if (e->graph) {
    hipGraphExecDestroy(e->graph);
    e->graph = nullptr;
    e->graphSubsteps = 0;
  }
}

// Switch a world to the full material pass set: bond and candidate rows for
// every slot (2 x 192 B per slot) and the slot -> position map.  Once on, it
// stays on; the next step re-records the graph.
void enable_materials(sn_engine* e) {
  if (e->gen) return;
  const size_t NC = std::max(e->opts.max_particles, 1u);
  Dev& D = e->D;
  D.bond = dalloc<uint32_t>(e, NC * BOND_CAP);
  D.bondCand = dalloc<uint32_t>(e, NC * BOND_CAP);
  D.posOf = dalloc<uint32_t>(e, NC);
  HIPCHECK(hipMemsetAsync(D.bond, 0, sizeof(uint32_t) * NC * BOND_CAP, e->stream));
  HIPCHECK(hipMemsetAsync(D.bondCand, 0, sizeof(uint32_t) * NC * BOND_CAP, e->stream));
  HIPCHECK(hipMemsetAsync(D.posOf, 0, sizeof(uint32_t) * NC, e->stream));
  D.gen = 1u;
  e->gen = true;
  drop_graph(e);
}

// Empty the bond rows of slots [first, first + count) of the partition at partBase.
void clear_bond_rows(sn_engine* e, uint32_t partBase, uint32_t first, uint32_t count) {
  if (!e->gen || count == 0) return;
  const size_t off = (size_t)(partBase + first) * BOND_CAP, len = (size_t)count * BOND_CAP;
  HIPCHECK(hipMemsetAsync(e->D.bond + off, 0, sizeof(uint32_t) * len, e->stream));
  HIPCHECK(hipMemsetAsync(e->D.bondCand + off, 0, sizeof(uint32_t) * len, e->stream));
}

// What a scene needs of the world: 1 if a fluid slot's row has friction or
// bonds (the material passes must run), 0 if not; -1 with e->err set when
// the scene cannot run here at all.
int scene_needs(sn_engine* e, const std::vector<int32_t>& mats, const sn_scene& sc) {
  bool need = false;
  for (uint32_t i = 0; i < sc.n; ++i) {
    const int32_t* w = sc.state8 + (size_t)i * 8;
    if ((uint32_t)w[5] & FLAG_RIGID) {
      fail(e, SN_E_UNSUPPORTED, "[aether] FLAG_RIGID: rigid bodies are not part of this build");
      return -1;
    }
    if (i < sc.n_fluid) {
      const int32_t* M = &mats[(uint32_t)w[4] * MAT_WORDS];
      need = need || M[M_FRIC] != 0 || M[M_BONDK] != 0;
    }
  }
  return need ? 1 : 0;
}

void enqueue_scan(sn_engine* e, uint32_t* count, uint32_t n, uint32_t* out, bool clear) {
  uint32_t nb = div_up(std::max(n, 1u), SCAN_BLOCK);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, e->stream, count, n, e->D.blockSums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, e->stream, e->D.blockSums, nb);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(256), 0, e->stream, count, n, e->D.blockSums, out, clear);
}

// One substep of every active scene, enqueued on the engine stream.
// Kernel classes timed by sn_profile_step (index into its outputs, + 1).
enum ProfClass {
  PC_PREDICT = 0, PC_SCAN, PC_SCATTER, PC_CANON, PC_NBR, PC_SOLVEA, PC_COLOUR, PC_XNC, PC_STH, PC_FINISH,
  PC_IMPULSE, PC_LISTS, PC_APPLY, PC_FRICTION, PC_MATERIAL, PC_COUNT
};

// Record an event before (begin) / after (end) one launch while profiling.
struct ProfScope {
  sn_engine* e;
  int cls;
  ProfScope(sn_engine* e_, int cls_) : e(e_), cls(cls_) {
    if (on()) record();
  }
  ~ProfScope() {
    if (on()) record();
  }
  bool on() const { return e->profiling && ((e->profMask >> cls) & 1u); }
  void record() {
    if (e->evUsed >= e->evPool.size()) {
      hipEvent_t ev;
      HIPCHECK(hipEventCreate(&ev));
      e->evPool.push_back(ev);
      e->evClass.push_back(0);
    }
    e->evClass[e->evUsed] = cls;
    HIPCHECK(hipEventRecord(e->evPool[e->evUsed], e->stream));
    e->evUsed++;
  }
};

void enqueue_substep(sn_engine* e) {
  Dev& D = e->D;
  hipStream_t st = e->stream;
  // scheduled stir events due at this substep (applied in arming order)
  for (uint32_t gi = 0; gi < e->maxEventsPerSubstep; ++gi) {
    ProfScope ps(e, PC_IMPULSE);
    hipLaunchKernelGGL(k_impulse_mark, dim3(e->actBlocks), dim3(256), 0, st, D, gi);
    enqueue_scan(e, D.hitFlag, D.N, D.hitScan, false);
    hipLaunchKernelGGL(k_impulse_apply, dim3(e->actBlocks), dim3(256), 0, st, D, gi);
  }
  if (e->maxEventsPerSubstep) {
    ProfScope ps(e, PC_IMPULSE);
    hipLaunchKernelGGL(k_impulse_advance, dim3(div_up(e->S, 64)), dim3(64), 0, st, D, e->dActive,
                       e->maxEventsPerSubstep);
  }
  {
    ProfScope ps(e, PC_PREDICT);
    hipLaunchKernelGGL(k_predict_count, dim3(e->actBlocks), dim3(256), 0, st, D);
  }
  {
    ProfScope ps(e, PC_SCAN);
    enqueue_scan(e, D.cellCount, D.C, D.cellStart, true);
  }
  {
    ProfScope ps(e, PC_SCATTER);
    hipLaunchKernelGGL(k_scatter, dim3(e->actBlocks), dim3(256), 0, st, D);
  }
  {
    ProfScope ps(e, PC_LISTS);
    hipLaunchKernelGGL(k_build_lists, dim3(div_up(std::max(std::max(D.C, D.nTiles), 1u), 256)), dim3(256), 0, st, D);
  }
  {
    ProfScope ps(e, PC_CANON);
    hipLaunchKernelGGL(k_canon_gather, dim3(e->actBlocks), dim3(256), 0, st, D);
    hipLaunchKernelGGL(k_canon_heavy_sort, dim3(e->canonBlocks), dim3(SORT_THREADS), 0, st, D);
    hipLaunchKernelGGL(k_canon_heavy_rank, dim3(e->canonBlocks), dim3(256), 0, st, D);
    hipLaunchKernelGGL(k_canon_heavy_place, dim3(e->canonBlocks), dim3(256), 0, st, D);
  }
  {
    ProfScope ps(e, PC_NBR);
    hipLaunchKernelGGL(k_block_box, dim3(e->actBlocks), dim3(256), 0, st, D);
    hipLaunchKernelGGL(k_heavy_order, dim3(e->canonBlocks), dim3(256), 0, st, D);
    if (D.dbgOn) hipLaunchKernelGGL(k_nbr_tile<false>, dim3(e->nbrBlocks), dim3(256), 0, st, D);
    else hipLaunchKernelGGL(k_nbr_tile<true>, dim3(e->nbrBlocks), dim3(256), 0, st, D);
  }
  const uint32_t iters = e->opts.iters;
  const bool gen = D.gen != 0u;   // the material passes run (friction, containment, bonds)
  for (uint32_t it = 0; it < iters; ++it) {
    if (it > 0) {
      ProfScope ps(e, PC_SOLVEA);
      hipLaunchKernelGGL(k_solveA, dim3(e->actBlocks), dim3(256), 0, st, D);
    }
    for (uint32_t c = 0; c < 4; ++c) {
      {
        ProfScope ps(e, PC_COLOUR);
        if (e->sZero) hipLaunchKernelGGL(k_colour<true>, dim3(e->colourBlocks), dim3(COLOUR_THREADS), 0, st, D, c);
        else hipLaunchKernelGGL(k_colour<false>, dim3(e->colourBlocks), dim3(COLOUR_THREADS), 0, st, D, c);
      }
      ProfScope ps(e, PC_APPLY);
      hipLaunchKernelGGL(k_apply, dim3(e->applyBlocks), dim3(256), 0, st, D, c,
                         (uint32_t)(!gen && it + 1 == iters));
    }
    if (gen) {
      ProfScope ps(e, PC_FRICTION);
      hipLaunchKernelGGL(k_friction, dim3(e->actBlocks * FR_L), dim3(256), 0, st, D);
      hipLaunchKernelGGL(k_apply_fric, dim3(e->actBlocks), dim3(256), 0, st, D);
    }
  }
  if (gen) {
    ProfScope ps(e, PC_MATERIAL);
    hipLaunchKernelGGL(k_contain, dim3(e->actBlocks), dim3(256), 0, st, D);
    hipLaunchKernelGGL(k_finalize, dim3(e->actBlocks), dim3(256), 0, st, D);
  }
  {
    ProfScope ps(e, PC_XNC);
    hipLaunchKernelGGL(k_xnc, dim3(e->actBlocks), dim3(256), 0, st, D);
  }
  {
    ProfScope ps(e, PC_STH);
    hipLaunchKernelGGL(k_st_thermal_hash, dim3(e->actBlocks), dim3(256), 0, st, D, (uint32_t)gen);
  }
  if (gen) {
    ProfScope ps(e, PC_MATERIAL);
    hipLaunchKernelGGL(k_contain_stress, dim3(e->actBlocks), dim3(256), 0, st, D);
    hipLaunchKernelGGL(k_bond_form1, dim3(e->actBlocks), dim3(256), 0, st, D);
    hipLaunchKernelGGL(k_bond_form2, dim3(e->actBlocks), dim3(256), 0, st, D);
    hipLaunchKernelGGL(k_hash_gen, dim3(e->actBlocks), dim3(256), 0, st, D);
  }
  {
    ProfScope ps(e, PC_FINISH);
    hipLaunchKernelGGL(k_hash_finish, dim3(e->S), dim3(64), 0, st, D, e->dActive);
  }
}

// Every neighbour tile of the active scenes: up to tileG consecutive cells of
// one row; k_build_lists keeps the occupied ones each substep.
void build_tiles(sn_engine* e) {
  auto& t = e->tiles;
  t.clear();
  for (uint32_t s = 0; s < e->S; ++s) {
    if (!e->active[s] || e->sp[s].n == 0) continue;
    const SceneP& P = e->sp[s];
    for (uint32_t cy = 0; cy < P.gridH; ++cy)
      for (uint32_t cx = 0; cx < P.gridW; cx += e->tileG)
        t.push_back(make_uint2(P.cellBase + cy * P.gridW + cx, std::min(e->tileG, P.gridW - cx) | (s << 8)));
  }
}

void upload_layout(sn_engine* e) {
  // sidOf: sorted position -> scene for the live range of each active scene
  std::vector<uint32_t> sid(e->N, INACTIVE);
  uint32_t before = 0;
  for (uint32_t s = 0; s < e->S; ++s) {
    SceneP& P = e->sp[s];
    P.posShift = (int32_t)P.partBase - (int32_t)before;
    if (e->active[s]) {
      for (uint32_t k = 0; k < P.n; ++k) sid[P.partBase + k] = s;
      before += P.n;
    }
  }
  // the 256-slot blocks holding live slots of active scenes, ascending
  std::vector<uint32_t> blocks;
  for (uint32_t s = 0; s < e->S; ++s) {
    const SceneP& P = e->sp[s];
    if (!e->active[s] || P.n == 0) continue;
    for (uint32_t b = P.partBase / 256u; b <= (P.partBase + P.n - 1u) / 256u; ++b)
      if (blocks.empty() || blocks.back() < b) blocks.push_back(b);
  }
  if (blocks.empty()) blocks.push_back(0u);   // one idle workgroup (it still clears the work-list counts)
  e->actBlocks = (uint32_t)blocks.size();
  HIPCHECK(hipMemcpyAsync(e->dBlockMap, blocks.data(), sizeof(uint32_t) * blocks.size(), hipMemcpyHostToDevice,
                          e->stream));
  // inactive scenes keep no cells: their counts stay zero
  HIPCHECK(hipMemcpyAsync((void*)e->D.sidOf, sid.data(), sizeof(uint32_t) * e->N, hipMemcpyHostToDevice, e->stream));
  HIPCHECK(hipMemcpyAsync((void*)e->D.sp, e->sp.data(), sizeof(SceneP) * e->S, hipMemcpyHostToDevice, e->stream));
  HIPCHECK(hipMemcpyAsync(e->dActive, e->active.data(), e->S, hipMemcpyHostToDevice, e->stream));
  // cell -> scene for the cells of active scenes
  std::vector<uint32_t> cs(std::max(e->C, 1u), INACTIVE);
  for (uint32_t s = 0; s < e->S; ++s) {
    if (!e->active[s]) continue;
    const SceneP& P = e->sp[s];
    for (uint32_t k = 0; k < P.cellTotal; ++k) cs[P.cellBase + k] = s;
  }
  HIPCHECK(hipMemcpyAsync((void*)e->D.cellScene, cs.data(), sizeof(uint32_t) * e->C, hipMemcpyHostToDevice,
                          e->stream));
  build_tiles(e);
  e->D.nTiles = (uint32_t)e->tiles.size();
  if (e->D.nTiles)
    HIPCHECK(hipMemcpyAsync((void*)e->D.tileAll, e->tiles.data(), sizeof(uint2) * e->D.nTiles, hipMemcpyHostToDevice,
                            e->stream));
  HIPCHECK(hipStreamSynchronize(e->stream));
}

int check_fault(sn_engine* e) {
  std::vector<SceneD> sd(e->S);
  HIPCHECK(hipMemcpyAsync(sd.data(), e->D.sd, sizeof(SceneD) * e->S, hipMemcpyDeviceToHost, e->stream));
  HIPCHECK(hipStreamSynchronize(e->stream));
  for (uint32_t s = 0; s < e->S; ++s) {
    if (sd[s].fault & FAULT_MATERIAL) {
      return fail(e, SN_E_UNSUPPORTED,
                  "[aether] scene " + std::to_string(s) +
                      " reached a material with friction, bonds or containment, which this build does not run");
    }
  }
  return SN_OK;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

// The operand bounds behind the binary64 pair forms (solveA_term, solveB_part,
// xnc_fluid, st_sum): |G| < 2^27 and |W| < 2^18 in the kernel tables, and in
// every material row |xsph|, |gCoh|, |gCur|, |adhesion| < 2^22 and
// |conductivity| < 2^26 (mat_mix of two such values stays below the same
// bound).  The frozen image and material table meet them; a world whose
// tables or materials do not runs every scene with the integer forms.
bool pair_bounds(const int32_t* luts, const int32_t* mats) {
  auto below = [](int32_t v, int32_t b) { return v > -b && v < b; };
  for (uint32_t i = 0; i < TN; ++i)
    if (!below(luts[TN + i], G_BOUND) || !below(luts[i], 1 << 18)) return false;
  for (uint32_t m = 0; m < MAT_MAX; ++m) {
    const int32_t* M = mats + (size_t)m * MAT_WORDS;
    if (!below(M[M_XSPH], 1 << 22) || !below(M[M_GCOH], 1 << 22) || !below(M[M_GCUR], 1 << 22) ||
        !below(M[M_ADH], 1 << 22) || !below(M[M_COND], 1 << 26))
      return false;
  }
  return true;
}

int sn_create(const sn_opts* opts, const int32_t* luts, const int32_t* mats, sn_engine** out) {
  if (!opts || !luts || !mats || !out) return SN_E_ARG;
  *out = nullptr;
  sn_engine* e = new sn_engine();
  e->opts = *opts;
  if (e->opts.max_scenes == 0) e->opts.max_scenes = 1;
  if (e->opts.iters == 0) e->opts.iters = 4;
  if (e->opts.stream_cap == 0) e->opts.stream_cap = 4096;
  if (e->opts.iters > 16) {
    int rc = fail(e, SN_E_ARG, "[aether] iters must be an integer in 1..16");
    *out = e;
    return rc;
  }
  if (e->opts.max_scenes > (1u << (32 - COL_SCENE_SHIFT))) {   // scene index field of the work-list items
    int rc = fail(e, SN_E_ARG, "[aether] maxScenes above " + std::to_string(1u << (32 - COL_SCENE_SHIFT)));
    *out = e;
    return rc;
  }
  try {
    HIPCHECK(hipSetDevice((int)e->opts.device));
    HIPCHECK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    const uint32_t NC = e->opts.max_particles, CC = e->opts.max_cells, SC = e->opts.max_scenes;
    Dev& D = e->D;
    D.dbgOn = e->opts.no_debug ? 0u : 1u;
    D.ps = dalloc<int4>(e, NC);
    D.pt = dalloc<int4>(e, NC);
    D.id = dalloc<uint32_t>(e, NC);
    D.st = dalloc<int2>(e, NC);
    D.psY = dalloc<int4>(e, NC);
    D.ptY = dalloc<int4>(e, NC);
    D.idY = dalloc<uint32_t>(e, NC);
    D.stY = dalloc<int2>(e, NC);
    D.sidOf = dalloc<uint32_t>(e, NC);
    D.pred0 = dalloc<int2>(e, NC);
    D.cell0 = dalloc<uint32_t>(e, NC);
    D.rank0 = dalloc<uint32_t>(e, NC);
    D.bucket = dalloc<uint32_t>(e, NC);
    D.bucketId = dalloc<uint32_t>(e, NC);
    D.cellCount = dalloc<uint32_t>(e, CC);
    D.cellStart = dalloc<uint32_t>(e, (size_t)CC + 1);
    D.blockSums = dalloc<uint32_t>(e, div_up(std::max(CC, NC), SCAN_BLOCK) + 1);
    D.pred = dalloc<int2>(e, NC);
    D.cellY = dalloc<uint32_t>(e, NC);
    D.info = dalloc<uint32_t>(e, NC);
    D.rho = dalloc<int32_t>(e, NC);
    D.lam = dalloc<int32_t>(e, NC);
    D.pl4 = dalloc<int4>(e, NC);
    D.vr4 = dalloc<int4>(e, NC);
    D.nr4 = dalloc<int4>(e, NC);
    D.dp = dalloc<int2>(e, NC);
    D.vtmp = dalloc<int2>(e, NC);
    D.rinv = dalloc<int32_t>(e, NC);
    D.nrm = dalloc<int2>(e, NC);
    D.heat = dalloc<int32_t>(e, NC);
    D.corr = dalloc<int32_t>(e, NC);
    D.nbr = dalloc<uint32_t>(e, (size_t)NC * MAXNBR);
    D.NS = NC;
    D.nbrN = dalloc<uint32_t>(e, NC);
    D.drop = dalloc<uint32_t>(e, NC);
    D.crank = dalloc<uint32_t>(e, NC);
    D.hperm = dalloc<uint32_t>(e, NC);
    // a binned cell holds more than TILE_CHUNK slots, so at most NC / 65 are
    // binned in one substep
    D.binCap = NC / 64u + 1u;
    D.binTab = dalloc<uint4>(e, (size_t)D.binCap * 64u);
    D.binBox = dalloc<int4>(e, (size_t)D.binCap * 64u);
    // every aligned 64-slot block of the 256-slot blocks; empty boxes until
    // first written (a sweep reads only blocks holding live candidates)
    {
      const size_t nb = (size_t)div_up(std::max(NC, 1u), 256) * 4u;
      std::vector<int4> empty(nb, make_int4(I32_MAX, I32_MAX, I32_MIN, I32_MIN));
      D.blkBox = dalloc<int4>(e, nb);
      HIPCHECK(hipMemcpy(D.blkBox, empty.data(), sizeof(int4) * nb, hipMemcpyHostToDevice));
    }
    e->dBlockMap = dalloc<uint32_t>(e, div_up(std::max(NC, 1u), 256));
    D.blockMap = e->dBlockMap;
    {
      const uint32_t zero = 0;   // a valid map of one block until the first layout upload
      HIPCHECK(hipMemcpy(e->dBlockMap, &zero, sizeof(zero), hipMemcpyHostToDevice));
    }
    HIPCHECK(hipMemset(D.crank, 0, sizeof(uint32_t) * std::max(NC, 1u)));
    D.hpart = dalloc<uint32_t>(e, (size_t)div_up(NC, 256) * HPART_WORDS);
    D.hitFlag = dalloc<uint32_t>(e, NC);
    D.hitScan = dalloc<uint32_t>(e, (size_t)NC + 1);
    e->dImpulse = dalloc<Event>(e, 1);
    int4* lut4 = dalloc<int4>(e, TN);
    int32_t* lutA = dalloc<int32_t>(e, TN);
    std::vector<int4> h4(TN);
    for (uint32_t i = 0; i < TN; ++i) h4[i] = make_int4(luts[i], luts[TN + i], luts[2 * TN + i], luts[4 * TN + i]);
    e->pairBounded = pair_bounds(luts, mats);
    e->sZero = true;
    for (uint32_t i = 0; i < TN; ++i) e->sZero = e->sZero && luts[4 * TN + i] == 0;
    HIPCHECK(hipMemcpy(lut4, h4.data(), sizeof(int4) * TN, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(lutA, luts + 3 * TN, sizeof(int32_t) * TN, hipMemcpyHostToDevice));
    D.lut4 = lut4;
    D.lutA = lutA;
    int32_t* dm = dalloc<int32_t>(e, MAT_MAX * MAT_WORDS);
    HIPCHECK(hipMemcpy(dm, mats, sizeof(int32_t) * MAT_MAX * MAT_WORDS, hipMemcpyHostToDevice));
    D.mats = dm;
    D.sealRows = 0;
    for (uint32_t m = 0; m < MAT_MAX; ++m) {
      const int32_t* M = mats + (size_t)m * MAT_WORDS;
      if (M[M_BONDK] != 0 && M[M_BONDREFORM] == 0 && M[M_SOLIDMODE] != 0 && M[M_PHANTOM] == 0)
        D.sealRows |= 1ull << m;   // seal_class (src/sim.wgsl:1978-1981)
    }
    D.sp = dalloc<SceneP>(e, SC);
    D.sd = dalloc<SceneD>(e, SC);
    D.hacc = dalloc<uint32_t>(e, (size_t)SC * 8);
    D.hring = dalloc<uint32_t>(e, (size_t)SC * e->opts.stream_cap * 8);
    D.ringCap = e->opts.stream_cap;
    D.events = dalloc<Event>(e, 1);
    e->dActive = dalloc<uint8_t>(e, SC);
    // per-substep work lists: at most one item per cell (tiles <= cells)
    D.cellScene = dalloc<uint32_t>(e, CC);
    D.tileAll = dalloc<uint2>(e, CC);
    D.cellBin = dalloc<uint32_t>(e, CC);
    // a list holds at most one item per cell (tile) plus one per full chunk
    // at most one item per cell or tile plus one per full 32 slots (every list's
    // chunks are of at least 32)
    for (uint32_t c = 0; c < 7; ++c) D.dlist[c] = dalloc<uint2>(e, (size_t)CC + NC / 32 + 64);
    D.listCount = dalloc<uint32_t>(e, 16);
    // persistent grids: as many workgroups as can be resident, a multiple of 8 (XCDs)
    {
      hipDeviceProp_t prop;
      HIPCHECK(hipGetDeviceProperties(&prop, (int)e->opts.device));
      int bN = 0, bC = 0, bA = 0;
      HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &bN, D.dbgOn ? reinterpret_cast<const void*>(k_nbr_tile<false>) : reinterpret_cast<const void*>(k_nbr_tile<true>),
          256, 0));
      HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bC, e->sZero ? reinterpret_cast<const void*>(k_colour<true>)
                                                                       : reinterpret_cast<const void*>(k_colour<false>),
                                                              COLOUR_THREADS, 0));
      HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bA, reinterpret_cast<const void*>(k_apply), 256, 0));
      const uint32_t cus = (uint32_t)std::max(prop.multiProcessorCount, 1);
      e->nbrBlocks = std::max(8u, (cus * (uint32_t)std::max(bN, 1)) & ~7u);
      e->colourBlocks = std::max(8u, (cus * (uint32_t)std::max(bC, 1)) & ~7u);
      e->applyBlocks = std::max(8u, (cus * (uint32_t)std::max(bA, 1)) & ~7u);
      e->canonBlocks = 2u * cus;
    }
    HIPCHECK(hipMemset(D.cellCount, 0, sizeof(uint32_t) * CC));
    if (e->opts.materials) {
      enable_materials(e);
      HIPCHECK(hipStreamSynchronize(e->stream));
    }
  } catch (const DevError& d) {
    *out = e;
    return fail(e, SN_E_DEVICE, d.msg);
  }
  *out = e;
  return SN_OK;
}

void sn_destroy(sn_engine* e) {
  if (!e) return;
  hipSetDevice((int)e->opts.device);
  if (e->stream) hipStreamSynchronize(e->stream);
  drop_graph(e);
  for (void* p : e->allocs) hipFree(p);
  if (e->dEvents) hipFree(e->dEvents);
  for (hipEvent_t ev : e->evPool) hipEventDestroy(ev);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
}

const char* sn_last_error(const sn_engine* e) { return e ? e->err.c_str() : "null engine"; }

int sn_load_scenes(sn_engine* e, const sn_scene* scenes, uint32_t count) {
  if (!e || (!scenes && count)) return SN_E_ARG;
  if (count == 0) return fail(e, SN_E_ARG, "[aether] a world needs at least one scene");
  if (count > e->opts.max_scenes)
    return fail(e, SN_E_CAPACITY, "[aether] " + std::to_string(count) + " scenes exceed maxScenes " +
                                      std::to_string(e->opts.max_scenes));
  uint64_t totN = 0, totC = 0;
  std::vector<SceneP> sp(count);
  for (uint32_t s = 0; s < count; ++s) {
    const sn_scene& sc = scenes[s];
    if (sc.n_fluid > sc.n) return fail(e, SN_E_ARG, "[aether] scene.nFluid out of range");
    if (sc.n && !sc.state8) return fail(e, SN_E_ARG, "[aether] scene state missing");
    if (sc.cell_shift < 1 || sc.cell_shift > 27)
      return fail(e, SN_E_ARG, "[aether] scene.cellShift outside the 1..27 contract");
    if ((int64_t(1) << sc.cell_shift) < H)
      return fail(e, SN_E_ARG, "[aether] cells smaller than the kernel radius");
    if (sc.dom_w <= 2 * WALL || sc.dom_h <= 2 * WALL)
      return fail(e, SN_E_ARG, "[aether] domain smaller than the wall band");
    uint64_t cs = uint64_t(1) << sc.cell_shift;
    uint32_t gw = (uint32_t)std::max<uint64_t>(1, ((uint64_t)sc.dom_w + cs - 1) / cs);
    uint32_t gh = (uint32_t)std::max<uint64_t>(1, ((uint64_t)sc.dom_h + cs - 1) / cs);
    if ((uint64_t)gw * gh > 1048576u)
      return fail(e, SN_E_ARG, "[aether] cellTotal exceeds the 2-level scan limit 1048576");
    bool bounded = sc.dom_w <= POS_BOUND && sc.dom_h <= POS_BOUND;
    for (uint32_t i = 0; i < sc.n; ++i) {
      const int32_t* w = sc.state8 + (size_t)i * 8;
      uint32_t m = (uint32_t)w[4];
      if (m >= MAT_MAX) return fail(e, SN_E_ARG, "[aether] matId outside the material table");
      bounded = bounded && w[0] >= -POS_BOUND && w[0] <= POS_BOUND && w[1] >= -POS_BOUND && w[1] <= POS_BOUND;
    }
    SceneP& P = sp[s];
    std::memset(&P, 0, sizeof(P));
    P.n = sc.n;
    P.nFluid = sc.n_fluid;
    P.gridW = gw;
    P.gridH = gh;
    P.cellShift = sc.cell_shift;
    P.cellTotal = gw * gh;
    P.cellBase = (uint32_t)totC;
    P.partBase = (uint32_t)totN;
    P.cap = sc.n;
    P.cellCap = gw * gh;
    P.domW = sc.dom_w;
    P.domH = sc.dom_h;
    P.vmax = 65536;
    P.weldHold = 0;
    P.bounded = (bounded && e->pairBounded) ? 1u : 0u;
    totN += sc.n;
    totC += (uint64_t)gw * gh;
  }
  if (totN > e->opts.max_particles)
    return fail(e, SN_E_CAPACITY, "[aether] scene.n = " + std::to_string(totN) + " exceeds maxParticles = " +
                                      std::to_string(e->opts.max_particles));
  if (totC > e->opts.max_cells)
    return fail(e, SN_E_CAPACITY, "[aether] cellTotal " + std::to_string(totC) + " exceeds allocated maxCells " +
                                      std::to_string(e->opts.max_cells));
  try {
    HIPCHECK(hipSetDevice((int)e->opts.device));
    HIPCHECK(hipStreamSynchronize(e->stream));
    drop_graph(e);
    // rows with friction or bonds switch the world to the material passes
    std::vector<int32_t> mats(MAT_MAX * MAT_WORDS);
    HIPCHECK(hipMemcpy(mats.data(), e->D.mats, sizeof(int32_t) * mats.size(), hipMemcpyDeviceToHost));
    bool need = e->gen;
    for (uint32_t s = 0; s < count; ++s) {
      const int r = scene_needs(e, mats, scenes[s]);
      if (r < 0) return SN_E_UNSUPPORTED;
      need = need || r > 0;
    }
    if (need)
      for (uint32_t s = 0; s < count; ++s)
        if (scenes[s].n > MAX_PARTICLES_BONDED)
          return fail(e, SN_E_CAPACITY, "[aether] scene.n = " + std::to_string(scenes[s].n) +
                                            " exceeds MAX_PARTICLES_BONDED = 131071 (bond partner field)");
    if (need) enable_materials(e);
    e->S = count;
    e->N = (uint32_t)totN;
    e->C = (uint32_t)totC;
    e->sp = sp;
    e->active.assign(count, 1);
    e->slotted = false;
    Dev& D = e->D;
    D.N = e->N;
    D.C = e->C;
    D.S = e->S;
    // state in slot order (the first substep sorts it)
    std::vector<int4> ps(e->N), pt(e->N);
    std::vector<uint32_t> id(e->N);
    std::vector<int2> st(e->N, make_int2(0, 0));
    for (uint32_t s = 0; s < count; ++s) {
      const sn_scene& sc = scenes[s];
      for (uint32_t i = 0; i < sc.n; ++i) {
        const int32_t* w = sc.state8 + (size_t)i * 8;
        uint32_t p = sp[s].partBase + i;
        ps[p] = make_int4(w[0], w[1], w[2], w[3]);
        pt[p] = make_int4(w[4], w[5], w[6], w[7]);
        id[p] = i;
        if (sc.st_acc) st[p] = make_int2(sc.st_acc[2 * i], sc.st_acc[2 * i + 1]);
      }
    }
    if (e->N) {
      HIPCHECK(hipMemcpyAsync(D.ps, ps.data(), sizeof(int4) * e->N, hipMemcpyHostToDevice, e->stream));
      HIPCHECK(hipMemcpyAsync(D.pt, pt.data(), sizeof(int4) * e->N, hipMemcpyHostToDevice, e->stream));
      HIPCHECK(hipMemcpyAsync(D.id, id.data(), sizeof(uint32_t) * e->N, hipMemcpyHostToDevice, e->stream));
      HIPCHECK(hipMemcpyAsync(D.st, st.data(), sizeof(int2) * e->N, hipMemcpyHostToDevice, e->stream));
    }
    std::vector<SceneD> sd(count);
    for (auto& d : sd) {
      std::memset(&d, 0, sizeof(d));
      d.bondDirty = 1;
      for (int j = 0; j < 8; ++j) d.chain[j] = CHAIN_INIT[j];
    }
    HIPCHECK(hipMemcpyAsync(D.sd, sd.data(), sizeof(SceneD) * count, hipMemcpyHostToDevice, e->stream));
    HIPCHECK(hipMemsetAsync(D.cellCount, 0, sizeof(uint32_t) * std::max(e->C, 1u), e->stream));
    for (uint32_t s = 0; s < count; ++s) clear_bond_rows(e, sp[s].partBase, 0, sp[s].cap);
    e->maxEventsPerSubstep = 0;
    // neighbour tile width from the initial occupancy of occupied cells
    {
      uint64_t occupied = 0, parts = 0;
      for (uint32_t s = 0; s < count; ++s) {
        const sn_scene& sc = scenes[s];
        std::vector<uint8_t> seen(sp[s].cellTotal, 0);
        for (uint32_t i = 0; i < sc.n; ++i) {
          int32_t x = sc.state8[(size_t)i * 8], y = sc.state8[(size_t)i * 8 + 1];
          int32_t cx = std::min(std::max(x >> sp[s].cellShift, 0), (int32_t)sp[s].gridW - 1);
          int32_t cy = std::min(std::max(y >> sp[s].cellShift, 0), (int32_t)sp[s].gridH - 1);
          uint8_t& v = seen[(size_t)cy * sp[s].gridW + cx];
          occupied += v ? 0 : 1;
          v = 1;
        }
        parts += sc.n;
      }
      double avg = std::max(occupied ? (double)parts / (double)occupied : 64.0, 1.0);
      // a tile serves up to 64 * NBR_PPL particles in one sweep
      uint32_t T = (uint32_t)(64.0 * NBR_PPL * 0.6 / avg);
      e->tileG = std::min(8u, std::max(1u, T));
    }
    upload_layout(e);
    e->loaded = true;
    e->weldSeq = 1;   // one loadScene is one bonding group; the sequence restarts (src/engine.js:3795-3801)
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn_load_scene(sn_engine* e, const sn_scene* scene) { return sn_load_scenes(e, scene, 1); }

// ------------------------------------------------------------------ slots
// A slotted world (continuous batching): `count` fixed partitions, each with
// room for cap_particles[s] slots and cap_cells[s] grid cells, all empty and
// inactive.  sn_load_slot then places one scene into a free partition while
// the others keep their state, and sn_set_active(…, 0) retires it.
int sn_reserve_slots(sn_engine* e, uint32_t count, const uint32_t* cap_particles, const uint32_t* cap_cells) {
  if (!e || !cap_particles || !cap_cells) return SN_E_ARG;
  if (count == 0) return fail(e, SN_E_ARG, "[aether] a world needs at least one slot");
  if (count > e->opts.max_scenes)
    return fail(e, SN_E_CAPACITY, "[aether] " + std::to_string(count) + " slots exceed maxScenes " +
                                      std::to_string(e->opts.max_scenes));
  uint64_t totN = 0, totC = 0;
  std::vector<SceneP> sp(count);
  for (uint32_t s = 0; s < count; ++s) {
    SceneP& P = sp[s];
    std::memset(&P, 0, sizeof(P));
    P.partBase = (uint32_t)totN;
    P.cellBase = (uint32_t)totC;
    P.cap = cap_particles[s];
    P.cellCap = cap_cells[s];
    P.vmax = 65536;
    P.domW = P.domH = 1;
    P.gridW = P.gridH = 1;
    P.cellShift = 18;
    totN += cap_particles[s];
    totC += cap_cells[s];
  }
  if (totN > e->opts.max_particles || totC > e->opts.max_cells)
    return fail(e, SN_E_CAPACITY, "[aether] slot capacities exceed maxParticles / maxCells");
  try {
    HIPCHECK(hipSetDevice((int)e->opts.device));
    HIPCHECK(hipStreamSynchronize(e->stream));
    drop_graph(e);
    e->S = count;
    e->N = (uint32_t)totN;
    e->C = (uint32_t)totC;
    e->sp = sp;
    e->active.assign(count, 0);
    Dev& D = e->D;
    D.N = e->N;
    D.C = e->C;
    D.S = e->S;
    std::vector<SceneD> sd(count);
    for (auto& d : sd) std::memset(&d, 0, sizeof(d));
    HIPCHECK(hipMemcpyAsync(D.sd, sd.data(), sizeof(SceneD) * count, hipMemcpyHostToDevice, e->stream));
    HIPCHECK(hipMemsetAsync(D.cellCount, 0, sizeof(uint32_t) * std::max(e->C, 1u), e->stream));
    for (uint32_t s = 0; s < count; ++s) clear_bond_rows(e, sp[s].partBase, 0, sp[s].cap);
    e->maxEventsPerSubstep = 0;
    e->tileG = 1;
    upload_layout(e);
    e->loaded = true;
    e->slotted = true;
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn_load_slot(sn_engine* e, uint32_t slot, const sn_scene* sc) {
  if (!e || !sc) return SN_E_ARG;
  if (!e->loaded || !e->slotted) return fail(e, SN_E_STATE, "[aether] sn_load_slot needs a slotted world");
  if (slot >= e->S) return fail(e, SN_E_ARG, "[aether] slot index out of range");
  if (sc->n_fluid > sc->n) return fail(e, SN_E_ARG, "[aether] scene.nFluid out of range");
  if (sc->n && !sc->state8) return fail(e, SN_E_ARG, "[aether] scene state missing");
  if (sc->cell_shift < 1 || sc->cell_shift > 27)
    return fail(e, SN_E_ARG, "[aether] scene.cellShift outside the 1..27 contract");
  if ((int64_t(1) << sc->cell_shift) < H) return fail(e, SN_E_ARG, "[aether] cells smaller than the kernel radius");
  if (sc->dom_w <= 2 * WALL || sc->dom_h <= 2 * WALL)
    return fail(e, SN_E_ARG, "[aether] domain smaller than the wall band");
  SceneP P = e->sp[slot];
  const uint64_t cs = uint64_t(1) << sc->cell_shift;
  const uint32_t gw = (uint32_t)std::max<uint64_t>(1, ((uint64_t)sc->dom_w + cs - 1) / cs);
  const uint32_t gh = (uint32_t)std::max<uint64_t>(1, ((uint64_t)sc->dom_h + cs - 1) / cs);
  if (sc->n > P.cap || (uint64_t)gw * gh > P.cellCap)
    return fail(e, SN_E_CAPACITY, "[aether] scene does not fit slot " + std::to_string(slot));
  bool bounded = sc->dom_w <= POS_BOUND && sc->dom_h <= POS_BOUND;
  for (uint32_t i = 0; i < sc->n; ++i) {
    const int32_t* w = sc->state8 + (size_t)i * 8;
    if ((uint32_t)w[4] >= MAT_MAX) return fail(e, SN_E_ARG, "[aether] matId outside the material table");
    bounded = bounded && w[0] >= -POS_BOUND && w[0] <= POS_BOUND && w[1] >= -POS_BOUND && w[1] <= POS_BOUND;
  }
  try {
    HIPCHECK(hipSetDevice((int)e->opts.device));
    HIPCHECK(hipStreamSynchronize(e->stream));
    std::vector<int32_t> mats(MAT_MAX * MAT_WORDS);
    HIPCHECK(hipMemcpy(mats.data(), e->D.mats, sizeof(int32_t) * mats.size(), hipMemcpyDeviceToHost));
    const int r = scene_needs(e, mats, *sc);
    if (r < 0) return SN_E_UNSUPPORTED;
    if ((r > 0 || e->gen) && sc->n > MAX_PARTICLES_BONDED)
      return fail(e, SN_E_CAPACITY, "[aether] scene.n = " + std::to_string(sc->n) +
                                        " exceeds MAX_PARTICLES_BONDED = 131071 (bond partner field)");
    if (r > 0) enable_materials(e);
    drop_graph(e);
    clear_bond_rows(e, P.partBase, 0, P.cap);
    P.n = sc->n;
    P.nFluid = sc->n_fluid;
    P.gridW = gw;
    P.gridH = gh;
    P.cellShift = sc->cell_shift;
    P.cellTotal = gw * gh;
    P.domW = sc->dom_w;
    P.domH = sc->dom_h;
    P.bounded = (bounded && e->pairBounded) ? 1u : 0u;
    e->sp[slot] = P;
    Dev& D = e->D;
    std::vector<int4> ps(sc->n), pt(sc->n);
    std::vector<uint32_t> id(sc->n);
    std::vector<int2> st(sc->n, make_int2(0, 0));
    for (uint32_t i = 0; i < sc->n; ++i) {
      const int32_t* w = sc->state8 + (size_t)i * 8;
      ps[i] = make_int4(w[0], w[1], w[2], w[3]);
      pt[i] = make_int4(w[4], w[5], w[6], w[7]);
      id[i] = i;
      if (sc->st_acc) st[i] = make_int2(sc->st_acc[2 * i], sc->st_acc[2 * i + 1]);
    }
    if (sc->n) {
      HIPCHECK(hipMemcpyAsync(D.ps + P.partBase, ps.data(), sizeof(int4) * sc->n, hipMemcpyHostToDevice, e->stream));
      HIPCHECK(hipMemcpyAsync(D.pt + P.partBase, pt.data(), sizeof(int4) * sc->n, hipMemcpyHostToDevice, e->stream));
      HIPCHECK(hipMemcpyAsync(D.id + P.partBase, id.data(), sizeof(uint32_t) * sc->n, hipMemcpyHostToDevice,
                              e->stream));
      HIPCHECK(hipMemcpyAsync(D.st + P.partBase, st.data(), sizeof(int2) * sc->n, hipMemcpyHostToDevice, e->stream));
    }
    SceneD d;
    std::memset(&d, 0, sizeof(d));
    d.bondDirty = 1;
    for (int j = 0; j < 8; ++j) d.chain[j] = CHAIN_INIT[j];
    HIPCHECK(hipMemcpyAsync(D.sd + slot, &d, sizeof(SceneD), hipMemcpyHostToDevice, e->stream));
    e->active[slot] = 1;
    upload_layout(e);
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn_step(sn_engine* e, uint32_t count) {
  if (!e) return SN_E_ARG;
  if (!e->loaded) return fail(e, SN_E_STATE, "[aether] step() called before loadScene()");
  if (count == 0) return SN_OK;
  try {
    HIPCHECK(hipSetDevice((int)e->opts.device));
    const uint32_t G = e->graphG;
    if (!e->graph) {
      hipGraph_t g;
      HIPCHECK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
      try {
        for (uint32_t k = 0; k < G; ++k) enqueue_substep(e);
      } catch (const DevError&) {
        hipGraph_t broken = nullptr;
        (void)hipStreamEndCapture(e->stream, &broken);   // leave the stream out of capture
        if (broken) (void)hipGraphDestroy(broken);
        throw;
      }
      HIPCHECK(hipStreamEndCapture(e->stream, &g));
      HIPCHECK(hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0));
      HIPCHECK(hipGraphDestroy(g));
      e->graphSubsteps = G;
    }
    uint32_t done = 0;
    while (count - done >= e->graphSubsteps) {
      HIPCHECK(hipGraphLaunch(e->graph, e->stream));
      done += e->graphSubsteps;
    }
    for (; done < count; ++done) enqueue_substep(e);
    HIPCHECK(hipGetLastError());
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

// The launch shape of a world, tuned on the world itself: every neighbour
// tile width (1, 2, 4, 8 cells) and hipGraph length (4, 8, 16 substeps) is
// timed over graph-replayed substeps.  A world's cost changes as its scenes
// evolve (blocks collapse and pack their cells), so the trials can start
// `lookahead` substeps ahead: the world is advanced that far once, and every
// trial starts from that state.  Before and after, the snapshot of what a
// substep carries to the next (state, stAcc, ids, the per-scene counters
// and cursors, bond rows) is restored, so tuning leaves no trace in any
// result; the fastest shape is kept.  Both knobs only regroup independent
// work: every list stays in canonical order and every sum is order-free.
int sn_autotune(sn_engine* e, uint32_t trial_substeps, uint32_t lookahead, sn_tune_trial* trials,
                uint32_t max_trials, uint32_t* n_trials, sn_tune_trial* chosen) {
  if (!e || !n_trials || !chosen || (max_trials && !trials)) return SN_E_ARG;
  if (!e->loaded) return fail(e, SN_E_STATE, "[aether] autotune() called before loadScene()");
  if (trial_substeps == 0) trial_substeps = 16;
  *n_trials = 0;
  std::vector<std::pair<void*, size_t>> regions;
  std::vector<void*> start, ahead;   // the world now, and `lookahead` substeps on
  hipEvent_t t0 = nullptr, t1 = nullptr;
  const uint32_t keepTile = e->tileG, keepG = e->graphG;
  int rc = SN_OK;
  auto put = [&](const std::vector<void*>& from) {
    for (size_t i = 0; i < from.size(); ++i)
      (void)hipMemcpyAsync(regions[i].first, from[i], regions[i].second, hipMemcpyDeviceToDevice, e->stream);
    (void)hipStreamSynchronize(e->stream);
  };
  try {
    HIPCHECK(hipSetDevice((int)e->opts.device));
    HIPCHECK(hipStreamSynchronize(e->stream));
    Dev& D = e->D;
    const size_t N = std::max(e->N, 1u);
    regions = {{D.ps, N * sizeof(int4)}, {D.pt, N * sizeof(int4)}, {D.id, N * sizeof(uint32_t)},
               {D.st, N * sizeof(int2)}, {D.sd, (size_t)e->S * sizeof(SceneD)}};
    if (e->gen) regions.push_back({D.bond, (size_t)std::max(e->opts.max_particles, 1u) * BOND_CAP * sizeof(uint32_t)});
    for (auto& r : regions) {
      void* p = nullptr;
      HIPCHECK(hipMalloc(&p, r.second));
      start.push_back(p);
      HIPCHECK(hipMemcpyAsync(p, r.first, r.second, hipMemcpyDeviceToDevice, e->stream));
    }
    if (lookahead) {
      rc = sn_step(e, lookahead);
      HIPCHECK(hipStreamSynchronize(e->stream));
      if (rc != SN_OK) throw DevError(e->err);
    }
    for (auto& r : regions) {
      void* p = nullptr;
      HIPCHECK(hipMalloc(&p, r.second));
      ahead.push_back(p);
      HIPCHECK(hipMemcpyAsync(p, r.first, r.second, hipMemcpyDeviceToDevice, e->stream));
    }
    HIPCHECK(hipEventCreate(&t0));
    HIPCHECK(hipEventCreate(&t1));
    const uint32_t tileW[4] = {1, 2, 4, 8}, graphLen[3] = {4, 8, 16};
    sn_tune_trial best{0, 0, 0.0f};
    for (uint32_t tw : tileW) {
      for (uint32_t gl : graphLen) {
        e->tileG = tw;
        e->graphG = gl;
        upload_layout(e);
        drop_graph(e);
        rc = sn_step(e, gl);   // capture and one warm replay
        const uint32_t steps = (trial_substeps + gl - 1) / gl * gl;
        if (rc == SN_OK) {
          HIPCHECK(hipEventRecord(t0, e->stream));
          rc = sn_step(e, steps);
          HIPCHECK(hipEventRecord(t1, e->stream));
        }
        HIPCHECK(hipStreamSynchronize(e->stream));
        put(ahead);
        if (rc != SN_OK) throw DevError(e->err);
        float ms = 0.0f;
        HIPCHECK(hipEventElapsedTime(&ms, t0, t1));
        const sn_tune_trial tr{tw, gl, ms / (float)steps};
        if (*n_trials < max_trials) trials[*n_trials] = tr;
        ++*n_trials;
        if (best.graph_substeps == 0 || tr.ms_per_substep < best.ms_per_substep) best = tr;
      }
    }
    e->tileG = best.tile_cells;
    e->graphG = best.graph_substeps;
    *chosen = best;
  } catch (const DevError& d) {
    e->tileG = keepTile;
    e->graphG = keepG;
    rc = fail(e, rc != SN_OK ? rc : SN_E_DEVICE, d.msg);
  }
  (void)hipStreamSynchronize(e->stream);
  put(start);   // the world as it was before the call, whatever happened
  try {
    upload_layout(e);
    drop_graph(e);
  } catch (const DevError& d) {
    if (rc == SN_OK) rc = fail(e, SN_E_DEVICE, d.msg);
  }
  for (void* p : start) hipFree(p);
  for (void* p : ahead) hipFree(p);
  if (t0) hipEventDestroy(t0);
  if (t1) hipEventDestroy(t1);
  return rc;
}

int sn_sync(sn_engine* e) {
  if (!e) return SN_E_ARG;
  try {
    HIPCHECK(hipSetDevice((int)e->opts.device));
    HIPCHECK(hipStreamSynchronize(e->stream));
    if (!e->loaded) return SN_OK;
    return check_fault(e);
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
}

int sn_query(sn_engine* e) {
  if (!e) return SN_E_ARG;
  hipError_t q = hipStreamQuery(e->stream);
  if (q == hipSuccess) return SN_OK;
  if (q == hipErrorNotReady) return 1;
  return fail(e, SN_E_DEVICE, std::string("hipStreamQuery: ") + hipGetErrorString(q));
}

int sn_profile_step(sn_engine* e, uint32_t count, uint32_t class_mask, float out_ms[SN_PROF_SLOTS],
                    uint32_t out_calls[SN_PROF_SLOTS]) {
  if (!e || !out_ms || !out_calls) return SN_E_ARG;
  if (!e->loaded) return fail(e, SN_E_STATE, "[aether] step() called before loadScene()");
  try {
    HIPCHECK(hipSetDevice((int)e->opts.device));
    for (int k = 0; k < SN_PROF_SLOTS; ++k) {
      out_ms[k] = 0.0f;
      out_calls[k] = 0;
    }
    if (count == 0) return SN_OK;
    e->profiling = true;
    e->profMask = class_mask;
    e->evUsed = 0;
    try {
      for (uint32_t k = 0; k < count; ++k) enqueue_substep(e);
    } catch (...) {
      e->profiling = false;
      throw;
    }
    e->profiling = false;
    const uint32_t used = e->evUsed;
    if (used < 2) {   // no launch of the selected classes
      HIPCHECK(hipStreamSynchronize(e->stream));
      out_calls[0] = count;
      return check_fault(e);
    }
    HIPCHECK(hipEventSynchronize(e->evPool[used - 1]));
    for (uint32_t i = 0; i + 1 < used; i += 2) {
      float t;
      HIPCHECK(hipEventElapsedTime(&t, e->evPool[i], e->evPool[i + 1]));
      out_ms[1 + e->evClass[i]] += t;
      out_calls[1 + e->evClass[i]] += 1;
    }
    float tot;
    HIPCHECK(hipEventElapsedTime(&tot, e->evPool[0], e->evPool[used - 1]));
    out_ms[0] = tot;
    out_calls[0] = count;
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return check_fault(e);
}

static int read_sd(sn_engine* e, uint32_t scene, SceneD& out) {
  if (!e->loaded) return fail(e, SN_E_STATE, "[aether] no scene loaded");
  if (scene >= e->S) return fail(e, SN_E_ARG, "[aether] scene index out of range");
  HIPCHECK(hipSetDevice((int)e->opts.device));
  HIPCHECK(hipStreamSynchronize(e->stream));
  HIPCHECK(hipMemcpy(&out, e->D.sd + scene, sizeof(SceneD), hipMemcpyDeviceToHost));
  if (out.fault & FAULT_MATERIAL) return check_fault(e);
  return SN_OK;
}

int sn_read_hash(sn_engine* e, uint32_t scene, uint32_t acc_out[8]) {
  if (!e || !acc_out) return SN_E_ARG;
  try {
    SceneD sd;
    int rc = read_sd(e, scene, sd);
    if (rc) return rc;
    Dev& D = e->D;
    HIPCHECK(hipMemsetAsync(D.hacc, 0, sizeof(uint32_t) * 8 * e->S, e->stream));
    hipLaunchKernelGGL(k_hash_state, dim3(div_up(std::max(e->sp[scene].n, 1u), 256)), dim3(256), 0, e->stream, D,
                       scene);
    HIPCHECK(hipMemcpyAsync(acc_out, D.hacc + 8 * scene, 32, hipMemcpyDeviceToHost, e->stream));
    HIPCHECK(hipStreamSynchronize(e->stream));
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn_read_hash_stream(sn_engine* e, uint32_t scene, uint32_t* out, size_t max_substeps, size_t* n_out) {
  if (!e || (!out && max_substeps) || !n_out) return SN_E_ARG;
  try {
    SceneD sd;
    int rc = read_sd(e, scene, sd);
    if (rc) return rc;
    size_t avail = sd.streamPos;
    if (avail > e->D.ringCap)
      return fail(e, SN_E_CAPACITY, "[aether] H ring overflowed: read the stream at least every " +
                                        std::to_string(e->D.ringCap) + " substeps");
    size_t n = std::min(avail, max_substeps);
    if (n)
      HIPCHECK(hipMemcpy(out, e->D.hring + (size_t)scene * e->D.ringCap * 8, n * 32, hipMemcpyDeviceToHost));
    *n_out = n;
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn_reset_stream(sn_engine* e) {
  if (!e) return SN_E_ARG;
  if (!e->loaded) return fail(e, SN_E_STATE, "[aether] no scene loaded");
  try {
    HIPCHECK(hipSetDevice((int)e->opts.device));
    std::vector<SceneD> sd(e->S);
    HIPCHECK(hipStreamSynchronize(e->stream));
    HIPCHECK(hipMemcpy(sd.data(), e->D.sd, sizeof(SceneD) * e->S, hipMemcpyDeviceToHost));
    for (auto& d : sd) d.streamPos = 0;
    HIPCHECK(hipMemcpy(e->D.sd, sd.data(), sizeof(SceneD) * e->S, hipMemcpyHostToDevice));
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn_read_chain(sn_engine* e, uint32_t scene, uint32_t chain_out[8]) {
  if (!e || !chain_out) return SN_E_ARG;
  try {
    SceneD sd;
    int rc = read_sd(e, scene, sd);
    if (rc) return rc;
    for (int j = 0; j < 8; ++j) chain_out[j] = sd.chain[j];
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn_debug_counters(sn_engine* e, uint32_t scene, uint32_t out[16]) {
  if (!e || !out) return SN_E_ARG;
  try {
    SceneD sd;
    int rc = read_sd(e, scene, sd);
    if (rc) return rc;
    for (int j = 0; j < 16; ++j) out[j] = sd.dbg[j];
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn_substep_count(sn_engine* e, uint32_t scene, uint32_t* out) {
  if (!e || !out) return SN_E_ARG;
  try {
    SceneD sd;
    int rc = read_sd(e, scene, sd);
    if (rc) return rc;
    *out = sd.substep;
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

static int read_slots(sn_engine* e, uint32_t scene, int32_t* st8, int32_t* dv16) {
  const SceneP& P = e->sp[scene];
  int32_t* dst = nullptr;
  int32_t* ddv = nullptr;
  HIPCHECK(hipMalloc((void**)&dst, std::max<size_t>(P.n, 1) * 32));
  if (dv16) HIPCHECK(hipMalloc((void**)&ddv, std::max<size_t>(P.n, 1) * 64));
  hipLaunchKernelGGL(k_to_slots, dim3(div_up(std::max(P.n, 1u), 256)), dim3(256), 0, e->stream, e->D, scene,
                     dst, ddv);
  if (st8) HIPCHECK(hipMemcpyAsync(st8, dst, (size_t)P.n * 32, hipMemcpyDeviceToHost, e->stream));
  if (dv16) HIPCHECK(hipMemcpyAsync(dv16, ddv, (size_t)P.n * 64, hipMemcpyDeviceToHost, e->stream));
  HIPCHECK(hipStreamSynchronize(e->stream));
  hipFree(dst);
  if (ddv) hipFree(ddv);
  return SN_OK;
}

int sn_read_state(sn_engine* e, uint32_t scene, int32_t* out, size_t words) {
  if (!e || !out) return SN_E_ARG;
  try {
    SceneD sd;
    int rc = read_sd(e, scene, sd);
    if (rc) return rc;
    if (words < (size_t)e->sp[scene].n * 8) return fail(e, SN_E_ARG, "[aether] readState buffer too small");
    return read_slots(e, scene, out, nullptr);
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
}

int sn_read_derived(sn_engine* e, uint32_t scene, int32_t* out, size_t words) {
  if (!e || !out) return SN_E_ARG;
  try {
    SceneD sd;
    int rc = read_sd(e, scene, sd);
    if (rc) return rc;
    if (words < (size_t)e->sp[scene].n * 16) return fail(e, SN_E_ARG, "[aether] readDerived buffer too small");
    return read_slots(e, scene, nullptr, out);
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
}

// restore(snap) with `derived` (src/demo_boot.js:308-321): after loadScene,
// the snapshot's state words and derived scratch are written back.  Of the
// derived words only stAcc (8..9) carries into the next substep (the
// others are recomputed by every substep), so the state and stAcc are
// rewritten in slot order, which the next substep sorts; the engine-side
// counters stay as they are.
int sn_write_state_derived(sn_engine* e, uint32_t scene, const int32_t* st8, size_t st_words, const int32_t* dv16,
                           size_t dv_words) {
  if (!e || !st8) return SN_E_ARG;
  if (!e->loaded) return fail(e, SN_E_STATE, "[aether] no scene loaded");
  if (scene >= e->S || !e->active[scene]) return fail(e, SN_E_ARG, "[aether] scene index out of range");
  const SceneP& P = e->sp[scene];
  if (st_words < (size_t)P.n * 8) return fail(e, SN_E_ARG, "[aether] state buffer too small");
  if (dv16 && dv_words < (size_t)P.n * 16) return fail(e, SN_E_ARG, "[aether] derived buffer too small");
  for (uint32_t i = 0; i < P.n; ++i)
    if ((uint32_t)st8[(size_t)i * 8 + 4] >= MAT_MAX) return fail(e, SN_E_ARG, "[aether] matId outside the material table");
  try {
    HIPCHECK(hipSetDevice((int)e->opts.device));
    HIPCHECK(hipStreamSynchronize(e->stream));
    std::vector<int4> ps(P.n), pt(P.n);
    std::vector<uint32_t> id(P.n);
    std::vector<int2> st(P.n, make_int2(0, 0));
    for (uint32_t i = 0; i < P.n; ++i) {
      const int32_t* w = st8 + (size_t)i * 8;
      ps[i] = make_int4(w[0], w[1], w[2], w[3]);
      pt[i] = make_int4(w[4], w[5], w[6], w[7]);
      id[i] = i;
      if (dv16) st[i] = make_int2(dv16[(size_t)i * 16 + 8], dv16[(size_t)i * 16 + 9]);
    }
    if (P.n) {
      Dev& D = e->D;
      HIPCHECK(hipMemcpy(D.ps + P.partBase, ps.data(), sizeof(int4) * P.n, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(D.pt + P.partBase, pt.data(), sizeof(int4) * P.n, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(D.id + P.partBase, id.data(), sizeof(uint32_t) * P.n, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(D.st + P.partBase, st.data(), sizeof(int2) * P.n, hipMemcpyHostToDevice));
    }
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn_read_bond(sn_engine* e, uint32_t scene, uint32_t* out, size_t words) {
  if (!e || !out) return SN_E_ARG;
  try {
    SceneD sd;
    int rc = read_sd(e, scene, sd);
    if (rc) return rc;
    const SceneP& P = e->sp[scene];
    const size_t need = (size_t)P.n * BOND_CAP;
    if (words < need) return fail(e, SN_E_ARG, "[aether] readBond buffer too small");
    if (!e->gen) {
      std::memset(out, 0, sizeof(uint32_t) * need);
      return SN_OK;
    }
    HIPCHECK(hipMemcpy(out, e->D.bond + (size_t)P.partBase * BOND_CAP, sizeof(uint32_t) * need,
                       hipMemcpyDeviceToHost));
    return SN_OK;
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
}

int sn_apply_impulse(sn_engine* e, uint32_t scene, int32_t cx, int32_t cy, int32_t r, int32_t ix, int32_t iy,
                     uint32_t limit, uint32_t* hits) {
  if (!e || !hits) return SN_E_ARG;
  // The event kernels read each scene's cursor into the schedule; for the
  // duration of this call every scene's cursor is pointed at a one-event
  // schedule (this scene) or at nothing (the others), and the saved cursors
  // are put back on every exit path, including a failed device call.
  std::vector<SceneD> saved;
  bool swapped = false;
  try {
    SceneD sd;
    int rc = read_sd(e, scene, sd);
    if (rc) return rc;
    Event ev{scene, sd.substep, cx, cy, r, ix, iy, limit};
    HIPCHECK(hipMemcpy(e->dImpulse, &ev, sizeof(Event), hipMemcpyHostToDevice));
    Dev D = e->D;
    D.events = e->dImpulse;
    saved.resize(e->S);
    HIPCHECK(hipMemcpy(saved.data(), e->D.sd, sizeof(SceneD) * e->S, hipMemcpyDeviceToHost));
    std::vector<SceneD> tmp = saved;
    for (uint32_t s = 0; s < e->S; ++s) {
      tmp[s].evBase = 0;
      tmp[s].evCursor = 0;
      tmp[s].evCount = (s == scene) ? 1u : 0u;
    }
    swapped = true;
    HIPCHECK(hipMemcpy(e->D.sd, tmp.data(), sizeof(SceneD) * e->S, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_impulse_mark, dim3(e->actBlocks), dim3(256), 0, e->stream, D, 0u);
    enqueue_scan(e, D.hitFlag, D.N, D.hitScan, false);
    hipLaunchKernelGGL(k_impulse_apply, dim3(e->actBlocks), dim3(256), 0, e->stream, D, 0u);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(e->stream));
    const SceneP& P = e->sp[scene];
    uint32_t total = 0;
    if (P.n) {
      uint32_t lastScan = 0, lastFlag = 0, base = 0;
      HIPCHECK(hipMemcpy(&base, D.hitScan + P.partBase, 4, hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(&lastScan, D.hitScan + P.partBase + P.n - 1, 4, hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(&lastFlag, D.hitFlag + P.partBase + P.n - 1, 4, hipMemcpyDeviceToHost));
      total = lastScan + lastFlag - base;
    }
    *hits = std::min(total, limit);
    saved[scene].pushHits += *hits;   // the saved records predate this call's count
    HIPCHECK(hipMemcpy(e->D.sd, saved.data(), sizeof(SceneD) * e->S, hipMemcpyHostToDevice));
    swapped = false;
  } catch (const DevError& d) {
    if (swapped) (void)hipMemcpy(e->D.sd, saved.data(), sizeof(SceneD) * e->S, hipMemcpyHostToDevice);
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn_read_push_hits(sn_engine* e, uint32_t scene, uint32_t* out) {
  if (!e || !out) return SN_E_ARG;
  try {
    SceneD sd;
    int rc = read_sd(e, scene, sd);
    if (rc) return rc;
    *out = sd.pushHits;
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn_set_events(sn_engine* e, const sn_event* events, uint32_t count) {
  if (!e || (!events && count)) return SN_E_ARG;
  if (!e->loaded) return fail(e, SN_E_STATE, "[aether] no scene loaded");
  try {
    HIPCHECK(hipSetDevice((int)e->opts.device));
    HIPCHECK(hipStreamSynchronize(e->stream));
    std::vector<SceneD> sd(e->S);
    HIPCHECK(hipMemcpy(sd.data(), e->D.sd, sizeof(SceneD) * e->S, hipMemcpyDeviceToHost));
    std::vector<Event> ev(count);
    std::vector<uint32_t> perSub;
    uint32_t maxGroup = 0;
    for (uint32_t k = 0; k < count; ++k) {
      const sn_event& x = events[k];
      if (x.scene >= e->S) return fail(e, SN_E_ARG, "[aether] event scene out of range");
      if (k && (x.scene < events[k - 1].scene ||
                (x.scene == events[k - 1].scene && x.substep < events[k - 1].substep)))
        return fail(e, SN_E_ARG, "[aether] events must be sorted by (scene, substep)");
      ev[k] = Event{x.scene, x.substep, x.cx, x.cy, x.r, x.ix, x.iy, x.limit};
    }
    for (auto& d : sd) { d.evBase = 0; d.evCount = 0; d.evCursor = 0; }
    for (uint32_t k = 0; k < count;) {
      uint32_t s = ev[k].scene, j = k;
      while (j < count && ev[j].scene == s) ++j;
      sd[s].evBase = k;
      sd[s].evCount = j - k;
      sd[s].evCursor = k;
      // skip events already in the past of this scene
      while (sd[s].evCursor < j && ev[sd[s].evCursor].substep < sd[s].substep) sd[s].evCursor++;
      for (uint32_t a = k; a < j;) {
        uint32_t b = a;
        while (b < j && ev[b].substep == ev[a].substep) ++b;
        maxGroup = std::max(maxGroup, b - a);
        a = b;
      }
      k = j;
    }
    if (count > e->eventCap) {
      if (e->dEvents) hipFree(e->dEvents);
      e->dEvents = nullptr;
      HIPCHECK(hipMalloc((void**)&e->dEvents, sizeof(Event) * count));
      e->eventCap = count;
    }
    if (count) HIPCHECK(hipMemcpy(e->dEvents, ev.data(), sizeof(Event) * count, hipMemcpyHostToDevice));
    e->D.events = e->dEvents;
    e->D.nEvents = count;
    HIPCHECK(hipMemcpy(e->D.sd, sd.data(), sizeof(SceneD) * e->S, hipMemcpyHostToDevice));
    e->maxEventsPerSubstep = maxGroup;
    drop_graph(e);
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn_set_active(sn_engine* e, const uint8_t* active, uint32_t count) {
  if (!e || !active) return SN_E_ARG;
  if (!e->loaded) return fail(e, SN_E_STATE, "[aether] no scene loaded");
  if (count != e->S) return fail(e, SN_E_ARG, "[aether] activity mask length != scene count");
  try {
    HIPCHECK(hipSetDevice((int)e->opts.device));
    HIPCHECK(hipStreamSynchronize(e->stream));
    // Deactivating a scene freezes it in its current order; the sorted layout
    // of each scene is self-contained, so the others are unaffected.
    for (uint32_t s = 0; s < count; ++s) e->active[s] = active[s] ? 1 : 0;
    upload_layout(e);
    drop_graph(e);
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn_spawn_fluid(sn_engine* e, const int32_t* fluid8, uint32_t k, uint32_t flags, uint32_t* n_out) {
  if (!e || !n_out || (k && !fluid8)) return SN_E_ARG;
  if (!e->loaded) return fail(e, SN_E_STATE, "[aether] no scene loaded");
  if (e->S != 1) return fail(e, SN_E_ARG, "[aether] spawnFluid needs a one-scene world");
  if (flags & ~(SN_SPAWN_NO_WELD | SN_SPAWN_OWN_GROUP)) return fail(e, SN_E_ARG, "[aether] spawnFluid: unknown option bits");
  const SceneP P = e->sp[0];
  *n_out = P.n;
  if (k == 0) return SN_OK;
  const uint32_t nb = P.n - P.nFluid;
  uint32_t newN = P.n + k;
  if (newN > e->opts.max_particles) return SN_OK;  // refused deterministically, n unchanged
  if (e->gen && newN > MAX_PARTICLES_BONDED) return SN_OK;
  try {
    SceneD sd;
    int rc = read_sd(e, 0, sd);
    if (rc) return rc;
    std::vector<int32_t> st((size_t)P.n * 8), dv((size_t)P.n * 16);
    rc = read_slots(e, 0, st.data(), dv.data());
    if (rc) return rc;
    std::vector<int32_t> mats(MAT_MAX * MAT_WORDS);
    HIPCHECK(hipMemcpy(mats.data(), e->D.mats, sizeof(int32_t) * mats.size(), hipMemcpyDeviceToHost));
    // One spawn event is one weld group: no gesture can be open in this
    // build, so every spawn takes the next id, with or without ownGroup
    // (src/engine.js:4570-4580,5933-5935).
    const uint32_t seq = (e->weldSeq % WELD_SEQ_MAX) + 1u;
    const uint32_t tag = (seq << 8) & WELD_TAG_MASK;
    std::vector<int32_t> nst((size_t)newN * 8), acc((size_t)newN * 2, 0);
    std::memcpy(nst.data(), st.data(), (size_t)P.nFluid * 32);
    for (uint32_t i = 0; i < k; ++i) {
      int32_t* o = nst.data() + (size_t)(P.nFluid + i) * 8;
      std::memcpy(o, fluid8 + (size_t)i * 8, 32);
      if ((uint32_t)o[4] >= MAT_MAX) return fail(e, SN_E_ARG, "[aether] spawnFluid: matId outside the material table");
      const int32_t* M = &mats[(size_t)(uint32_t)o[4] * MAT_WORDS];
      if (!(flags & SN_SPAWN_NO_WELD) && M[M_BONDGROUP] >= 0 && M[M_BONDREFORM] == 0)
        o[5] = (int32_t)((uint32_t)o[5] | tag);
    }
    // the boundary block moves up by k slots, rewritten as the reference
    // rewrites it from the loaded scene: position, material, temperature and
    // fuel kept (boundary slots never move and the thermal passes skip
    // them), velocity and flags zero (src/engine.js:5964-5971)
    for (uint32_t b = 0; b < nb; ++b) {
      const int32_t* src = st.data() + (size_t)(P.nFluid + b) * 8;
      int32_t* o = nst.data() + (size_t)(P.nFluid + k + b) * 8;
      o[0] = src[0]; o[1] = src[1]; o[2] = 0; o[3] = 0;
      o[4] = src[4]; o[5] = 0; o[6] = src[6]; o[7] = src[7];
    }
    for (uint32_t i = 0; i < P.nFluid; ++i) {
      acc[2 * i] = dv[(size_t)i * 16 + 8];
      acc[2 * i + 1] = dv[(size_t)i * 16 + 9];
    }
    // fluid slots keep their ids, so their bond rows carry over unchanged;
    // the new fluid slots start with empty rows (src/engine.js:5987-5992)
    std::vector<uint32_t> rows;
    if (e->gen && P.nFluid) {
      rows.resize((size_t)P.nFluid * BOND_CAP);
      HIPCHECK(hipMemcpy(rows.data(), e->D.bond + (size_t)P.partBase * BOND_CAP, sizeof(uint32_t) * rows.size(),
                         hipMemcpyDeviceToHost));
    }
    sn_scene sc{newN, P.nFluid + k, nst.data(), P.domW, P.domH, P.cellShift, acc.data()};
    rc = sn_load_scenes(e, &sc, 1);
    if (rc) return rc;
    e->weldSeq = seq;
    if (!rows.empty())
      HIPCHECK(hipMemcpy(e->D.bond + (size_t)e->sp[0].partBase * BOND_CAP, rows.data(),
                         sizeof(uint32_t) * rows.size(), hipMemcpyHostToDevice));
    // carry the engine-side counters across the reload; formation is dirty
    SceneD nd = sd;
    nd.bondDirty = 1;
    nd.evBase = nd.evCount = nd.evCursor = 0;
    for (int j = 0; j < 8; ++j) nd.acc[j] = 0;
    HIPCHECK(hipMemcpy(e->D.sd, &nd, sizeof(SceneD), hipMemcpyHostToDevice));
    *n_out = newN;
  } catch (const DevError& d) {
    return fail(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

}  // extern "C"

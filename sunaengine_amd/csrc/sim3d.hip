// MI355X-native deterministic 3D PBF substep — kernels and C ABI (sn3_*).
//
// Same method as the reference's 3D solver (3d/sim3d.wgsl, schedule
// 3d/app.js:673-697): predict -> grid -> canonical order -> 27-cell
// neighbour list -> (density, position correction) x iters -> integrate.
// As in the 2D engine, particles live permanently in (cell, slot id) order:
// the candidates of the canonical stencil order (dz, dy, dx, ascending id)
// are nine contiguous row ranges, and the neighbour list stores sorted
// positions in exactly that order.  That order matters here: the position
// correction is a SATURATING running sum (3d/sim3d.wgsl:511-514), which is
// folded neighbour by neighbour in list order, exactly as the reference does.
//
// Kernels per substep (iters = k):
//   predict_count   v and pred (3d/sim3d.wgsl:188-215), cell, arrival rank
//   scan x3         exclusive prefix of cell counts (any grid size)
//   scatter, canon_gather  (3d/sim3d.wgsl:304-334)
//   nbr_density     buildNbr fused with iteration 0's density (:341-468)
//   solve_apply     solveDp + applyDp (:476-537) into the other pred buffer;
//                   the last one also integrates (:542-552) and hashes
//   density         iterations 1..k-1
//   hash_finish     chain fold, H ring append
//
// H_set / H_slot for 3D: the reference's only 3D digest is a striped FNV
// over p.xyz, v.xyz (3d/sim3d.wgsl:571-590; also provided, sn3_fnv_digest).
// The per-substep hash here is digestState (src/hash/statehash.js:80-159)
// over each particle's words p.xyzw (w = material), v.xyzw with its slot id.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sunaengine.h"
#include "fixed.hpp"
#include "scan.hpp"

using namespace sn;

namespace {

constexpr uint32_t MAXNBR = 48;
constexpr int32_t H = 131072;
constexpr int32_t H2Q = 262144;
constexpr uint32_t OFF_W = 0, OFF_G = 2049, OFF_P = 4098, OFF_PG = 4611, OFF_BW = 5124, OFF_BG = 6148;
constexpr int32_t BALL_R2MAX = 4194304;
constexpr uint32_t OFF_MAT = 7172, MAT_WORDS = 4, NMAT3 = 6;
constexpr int32_t RHO0 = 1048576, C_MAX = 1048576, LAM_MAX = 4194304, EPS_S = 524288;
constexpr int32_t S_TERM = 8388608;
constexpr int32_t MARGIN = 32768;
constexpr uint32_t INACTIVE = 0xffffffffu;
constexpr uint32_t LUT3_WORDS = OFF_MAT + NMAT3 * MAT_WORDS;
constexpr uint32_t LUT_N3 = 2049;          // W and G bins
constexpr uint32_t MIXED = 0xfffffffeu;    // a wave/block spanning scenes
constexpr int HPART3 = 9;                  // 8 hash words + truncation count
constexpr uint32_t CHUNK3 = 64;            // particles per neighbour-sweep item

struct DevError3 {
  std::string msg;
  explicit DevError3(std::string m) : msg(std::move(m)) {}
};
#define HIPCHECK3(x)                                                             \
  do {                                                                           \
    hipError_t err_ = (x);                                                       \
    if (err_ != hipSuccess) throw DevError3(std::string(#x) + ": " + hipGetErrorString(err_)); \
  } while (0)

struct Scene3P {
  uint32_t n, gx, gy, gz;
  uint32_t cellBase, cellTotal, partBase;
  int32_t posShift;
  int32_t wx, wy, wz, pad;   // world extents in ticks
  uint32_t cap, cellCap, pad2, pad3;  // partition sizes (slotted worlds)
};
struct Scene3D {
  uint32_t substep, fault, streamPos, pad;
  uint32_t acc[8];
  uint32_t chain[8];
  uint32_t dbg[16];
  uint32_t evCursor, evCount, evBase, pad2;
};
struct Inputs3 {
  int32_t stirAmp, stirPhase;
  uint32_t ballOn;
  int32_t ballX, ballY, ballZ, ballReach;
};
struct Event3 {
  uint32_t scene, substep;
  int32_t cx, cy, cz, r, ix, iy, iz;
  uint32_t limit;
};

struct Dev3 {
  int4* ps;      // p.xyz, material   (current order)
  int4* vel;     // v.xyz, 0
  uint32_t* id;
  int4* psY;     // this substep's input snapshot, new order
  uint32_t* idY;
  uint32_t* matY;
  const uint32_t* sidOf;
  int4* pr0;     // pred, old order
  uint32_t* cell0;
  uint32_t* rank0;
  uint32_t* bucket;
  uint32_t* bucketId;
  uint32_t* cellCount;
  uint32_t* cellStart;
  uint32_t* blockSums;
  uint32_t* cellY;
  int4* prA;     // pred.xyz, lambda (ping)
  int4* prB;     // (pong)
  int4* aux;     // last dp.xyz, rho (for read-back of the full record)
  uint32_t* nbr;
  uint32_t* nbrN;
  uint32_t* drop;            // candidates dropped by truncation, per particle
  uint32_t* hpart;           // per-block partial sums [block][HPART3]
  const uint32_t* cellScene; // global cell -> scene, INACTIVE if its scene is inactive
  uint2* dlist;              // this substep's occupied cell chunks {cell, chunk}
  uint32_t* listCount;       // [1] entries of dlist (cleared by predict)
  uint32_t dbgOn;            // 0: the reference's debug:false build, counters stay zero
  const int32_t* luts;
  const Scene3P* sp;
  Scene3D* sd;
  uint32_t* hring;
  uint32_t* hacc;
  Inputs3 in;
  const Event3* events;
  uint32_t* hitFlag;
  uint32_t* hitScan;
  uint32_t N, C, S, ringCap;
  uint32_t NS;  // neighbour-list stride: nbr[k * NS + p]
};

__device__ __forceinline__ int32_t mat_grav(const Dev3& D, uint32_t m) { return D.luts[OFF_MAT + m * MAT_WORDS]; }
__device__ __forceinline__ int32_t mat_rhoW(const Dev3& D, uint32_t m) { return D.luts[OFF_MAT + m * MAT_WORDS + 1]; }
__device__ __forceinline__ int32_t mat_damp(const Dev3& D, uint32_t m) { return D.luts[OFF_MAT + m * MAT_WORDS + 2]; }

__device__ __forceinline__ int3 clamp_world(const Scene3P& P, int3 p) {
  return make_int3(clampi(p.x, MARGIN, P.wx - MARGIN), clampi(p.y, MARGIN, P.wy - MARGIN),
                   clampi(p.z, MARGIN, P.wz - MARGIN));
}
__device__ __forceinline__ uint32_t cell_of(const Scene3P& P, int3 p) {
  // p / H with p inside the world (clamp_world keeps it positive)
  uint32_t cx = (uint32_t)(p.x / H), cy = (uint32_t)(p.y / H), cz = (uint32_t)(p.z / H);
  return P.cellBase + (cz * P.gy + cy) * P.gx + cx;
}
__device__ __forceinline__ int32_t r2q_of(int3 d) {
  int32_t qx = d.x >> 8, qy = d.y >> 8, qz = d.z >> 8;
  return wadd(wadd(wmul(qx, qx), wmul(qy, qy)), wmul(qz, qz));
}
__device__ __forceinline__ int32_t s_term(int32_t a, Dbg& g) { return clampi(fp_mul_s(a, a, 16, g), 0, S_TERM); }

// wall_of (3d/sim3d.wgsl:129-141): distance to wall w and the direction
// toward it scaled by gm.
__device__ __forceinline__ int32_t wall_dist(const Scene3P& P, int3 p, uint32_t w) {
  uint32_t axis = w >> 1;
  int32_t pos = axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
  int32_t wmax = axis == 0 ? P.wx : (axis == 1 ? P.wy : P.wz);
  int32_t d = (w & 1u) ? wsub(wmax, pos) : pos;
  return maxi(d, 0);
}
__device__ __forceinline__ int3 wall_dir(uint32_t w, int32_t gm, Dbg& g) {
  uint32_t axis = w >> 1;
  int32_t s = (w & 1u) ? gm : fp_neg_sat(gm, g);
  return make_int3(axis == 0 ? s : 0, axis == 1 ? s : 0, axis == 2 ? s : 0);
}
// ball_term (3d/sim3d.wgsl:152-173)
__device__ __forceinline__ bool ball_term(const Dev3& D, int3 pi, int32_t& w, int3& gv, Dbg& g) {
  if (D.in.ballOn != 1u) return false;
  int3 dc = make_int3(wsub(D.in.ballX, pi.x), wsub(D.in.ballY, pi.y), wsub(D.in.ballZ, pi.z));
  if (fp_abs(dc.x, g) >= D.in.ballReach || fp_abs(dc.y, g) >= D.in.ballReach || fp_abs(dc.z, g) >= D.in.ballReach)
    return false;
  int32_t r2q = r2q_of(dc);
  if (r2q >= BALL_R2MAX) return false;
  uint32_t bi = (uint32_t)r2q >> 12;
  int32_t gt = D.luts[OFF_BG + bi];
  w = D.luts[OFF_BW + bi];
  gv = make_int3(fp_mul_s(gt, dc.x, 16, g), fp_mul_s(gt, dc.y, 16, g), fp_mul_s(gt, dc.z, 16, g));
  return true;
}

// Wrapping u32 sum over the 64 lanes (every lane must be active).
__device__ __forceinline__ uint32_t wave_sum3(uint32_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
  return v;
}
// Debug counters: summed per wave, one atomic per nonzero counter when the
// wave is uniform in scene.  Must be called with every lane active.
__device__ __forceinline__ void flush_dbg3(const Dev3& D, uint32_t s, const Dbg& g, uint32_t trunc) {
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
      x = wave_sum3(x);
      if ((threadIdx.x & 63u) == 0u && s0 != INACTIVE) atomicAdd(&D.sd[s0].dbg[slot[k]], x);
    } else if (s != INACTIVE && x) {
      atomicAdd(&D.sd[s].dbg[slot[k]], x);
    }
  }
}

// ------------------------------------------------------------------ predict
__device__ __forceinline__ uint32_t predict3(const Dev3& D, uint32_t p, uint32_t s, Dbg& g) {
  const Scene3P& P = D.sp[s];
  int4 a = D.ps[p];
  int4 v4 = D.vel[p];
  uint32_t mi = (uint32_t)a.w;
  int32_t vx = v4.x, vy = fp_sub_sat(v4.y, mat_grav(D, mi), g), vz = v4.z;
  if (D.in.stirAmp != 0) {
    int32_t tri = abs(((D.in.stirPhase + 32768) & 131071) - 65536) - 32768;
    vx = fp_add_sat(vx, fp_mul_s(tri, D.in.stirAmp, 16, g), g);
  }
  int32_t dmp = mat_damp(D, mi);
  vx = fp_mul_s(vx, dmp, 16, g);
  vy = fp_mul_s(vy, dmp, 16, g);
  vz = fp_mul_s(vz, dmp, 16, g);
  int3 pr = clamp_world(P, make_int3(fp_add_sat(a.x, vx, g), fp_add_sat(a.y, vy, g), fp_add_sat(a.z, vz, g)));
  uint32_t c = cell_of(P, pr);
  D.pr0[p] = make_int4(pr.x, pr.y, pr.z, 0);
  D.cell0[p] = c;
  return c;
}
__global__ __launch_bounds__(256) void k3_predict_count(Dev3 D) {
  uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) D.listCount[0] = 0u;  // rebuilt after the scan
  uint32_t s = p < D.N ? D.sidOf[p] : INACTIVE;
  Dbg g;
  const uint32_t c = s != INACTIVE ? predict3(D, p, s, g) : INACTIVE;
  // arrival rank in the cell: one atomic per distinct cell of the workgroup
  // instead of one per particle (any distinct ranks are valid: canonicalize
  // re-ranks by slot id)
  const uint32_t r = wg_cell_rank(D.cellCount, c, INACTIVE);
  if (c != INACTIVE) D.rank0[p] = r;
  flush_dbg3(D, s, g, 0);
}

__global__ __launch_bounds__(256) void k3_scatter(Dev3 D) {
  uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p >= D.N) return;
  uint32_t s = D.sidOf[p];
  if (s == INACTIVE) return;
  uint32_t q = (uint32_t)((int32_t)D.cellStart[D.cell0[p]] + D.sp[s].posShift) + D.rank0[p];
  D.bucket[q] = p;
  D.bucketId[q] = D.id[p];
}

__global__ __launch_bounds__(256) void k3_canon_gather(Dev3 D) {
  uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= D.N) return;
  uint32_t s = D.sidOf[t];
  if (s == INACTIVE) return;
  int32_t sh = D.sp[s].posShift;
  uint32_t me = D.bucket[t], myid = D.bucketId[t];
  uint32_t c = D.cell0[me];
  uint32_t lo = (uint32_t)((int32_t)D.cellStart[c] + sh), hi = (uint32_t)((int32_t)D.cellStart[c + 1] + sh);
  uint32_t rank = 0;
  for (uint32_t u = lo; u < hi; ++u) rank += (D.bucketId[u] < myid) ? 1u : 0u;
  uint32_t q = lo + rank;
  int4 a = D.ps[me];
  D.psY[q] = a;
  D.idY[q] = myid;
  D.matY[q] = (uint32_t)a.w;
  D.prA[q] = D.pr0[me];
  D.cellY[q] = c;
}

// ------------------------------------------------------------------ density
struct DensAcc {
  int32_t rho, S, sx, sy, sz;
};
__device__ __forceinline__ void dens_term(const Dev3& D, DensAcc& A, int3 d, uint32_t ui, uint32_t mj, Dbg& g) {
  A.rho = wadd(A.rho, wmul(D.luts[OFF_W + ui], wadd(256, mat_rhoW(D, mj))) >> 8);
  int32_t gt = D.luts[OFF_G + ui];
  int32_t gx = fp_mul_s(gt, d.x, 16, g), gy = fp_mul_s(gt, d.y, 16, g), gz = fp_mul_s(gt, d.z, 16, g);
  A.sx = wadd(A.sx, gx);
  A.sy = wadd(A.sy, gy);
  A.sz = wadd(A.sz, gz);
  A.S = wadd(wadd(wadd(A.S, s_term(gx, g)), s_term(gy, g)), s_term(gz, g));
}
__device__ __forceinline__ void dens_finish(const Dev3& D, const Scene3P& P, uint32_t p, int3 pi, DensAcc& A,
                                            int4* prOut, Dbg& g) {
  for (uint32_t w = 0; w < 6; ++w) {
    int32_t dist = wall_dist(P, pi, w);
    if (dist >= H) continue;
    uint32_t b = (uint32_t)dist >> 8;
    A.rho = wadd(A.rho, D.luts[OFF_P + b]);
    int3 gv = wall_dir(w, D.luts[OFF_PG + b], g);
    A.sx = wadd(A.sx, gv.x);
    A.sy = wadd(A.sy, gv.y);
    A.sz = wadd(A.sz, gv.z);
    A.S = wadd(wadd(wadd(A.S, s_term(gv.x, g)), s_term(gv.y, g)), s_term(gv.z, g));
  }
  int32_t bw;
  int3 bg;
  if (ball_term(D, pi, bw, bg, g)) {
    A.rho = wadd(A.rho, bw);
    A.sx = wadd(A.sx, bg.x);
    A.sy = wadd(A.sy, bg.y);
    A.sz = wadd(A.sz, bg.z);
    A.S = wadd(wadd(wadd(A.S, s_term(bg.x, g)), s_term(bg.y, g)), s_term(bg.z, g));
  }
  A.S = wadd(wadd(wadd(A.S, s_term(A.sx, g)), s_term(A.sy, g)), s_term(A.sz, g));
  int32_t Cq = clampi(fp_shr_rne(wsub(A.rho, RHO0), 4), 0, C_MAX);
  int32_t lam = clampi(fp_neg_sat(fp_divshift(Cq, wadd(A.S, EPS_S), 24, g), g), fp_neg_sat(LAM_MAX, g), LAM_MAX);
  prOut[p].w = lam;
  D.aux[p].w = A.rho;
}

// ============================================================ per-substep work
// After the scan: one item {cell, chunk} per CHUNK3 particles of every
// occupied cell of an active scene, appended with one atomic per workgroup.
// Items are independent, so list order (left open by the atomics) affects
// no result; chunking bounds each item's work however densely a cell packs.
// Called with every lane of every wave active.
__global__ __launch_bounds__(256) void k3_build_lists(Dev3 D) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  uint32_t nch = 0;
  if (t < D.C && D.cellScene[t] != INACTIVE) nch = (D.cellStart[t + 1] - D.cellStart[t] + CHUNK3 - 1u) / CHUNK3;
  wg_append_chunks(nch, &D.listCount[0], D.dlist, t);
}

// XCD x (= blockIdx.x % 8) walks the x-th eighth of a device list.
struct ListRange3 {
  uint32_t lo, hi, first, stride;
};
__device__ __forceinline__ ListRange3 xcd_range3(uint32_t n, uint32_t unitsPerBlock, uint32_t unit) {
  const uint32_t xcd = blockIdx.x & 7u, nb = gridDim.x >> 3, lb = blockIdx.x >> 3;
  ListRange3 r;
  r.lo = (uint32_t)(((uint64_t)n * xcd) >> 3);
  r.hi = (uint32_t)(((uint64_t)n * (xcd + 1u)) >> 3);
  r.first = r.lo + lb * unitsPerBlock + unit;
  r.stride = nb * unitsPerBlock;
  return r;
}
__device__ __forceinline__ int32_t wave_min3(int32_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = mini(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ int32_t wave_max3(int32_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = maxi(v, __shfl_xor(v, off));
  return v;
}

// buildNbr (3d/sim3d.wgsl:341-390) fused with iteration 0's density.  One
// wave per chunk of up to 64 particles of one cell, one particle per lane.
// The chunk's candidates are the cell's 27-cell stencil: nine (dz, dy) rows,
// each a contiguous sorted range over x in cx-1..cx+1, visited in exactly the
// canonical order (dz, dy, dx, ascending id).  They are loaded 64 at a time,
// one per lane, and broadcast with v_readlane.  A candidate is accepted only
// with |d| < H on every axis; one outside the chunk's bounding box grown by
// H - 1 is therefore rejected by every lane (positions lie inside the world,
// so no difference wraps), and each block is reduced to a ballot of the
// candidates inside that box.  The kernel tables (W, G) sit in LDS.
__global__ __launch_bounds__(256) void k3_nbr_tile(Dev3 D) {
  __shared__ int32_t lutW[LUT_N3], lutG[LUT_N3];
  __shared__ int32_t rhoW[NMAT3];
  __shared__ int4 stage[4][64];   // a block's candidates in reach, per wave
  for (uint32_t i = threadIdx.x; i < LUT_N3; i += 256u) {
    lutW[i] = D.luts[OFF_W + i];
    lutG[i] = D.luts[OFF_G + i];
  }
  if (threadIdx.x < NMAT3) rhoW[threadIdx.x] = mat_rhoW(D, threadIdx.x);
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u;
  int4* stg = stage[threadIdx.x >> 6];
  const uint32_t nwork = (uint32_t)__builtin_amdgcn_readfirstlane((int)D.listCount[0]);
  const ListRange3 R = xcd_range3(nwork, 4u, threadIdx.x >> 6);
  for (uint32_t w = R.first; w < R.hi; w += R.stride) {
    const uint2 item = D.dlist[w];
    const uint32_t c = __builtin_amdgcn_readfirstlane(item.x);
    const uint32_t chunk = __builtin_amdgcn_readfirstlane(item.y);
    const uint32_t s = __builtin_amdgcn_readfirstlane(D.cellScene[c]);
    const Scene3P& P = D.sp[s];
    const int32_t sh = P.posShift;
    const uint32_t plo = (uint32_t)((int32_t)D.cellStart[c] + sh) + chunk * CHUNK3;
    const uint32_t phi = (uint32_t)mini((int32_t)((int32_t)D.cellStart[c + 1] + sh), (int32_t)(plo + CHUNK3));
    const uint32_t lc = c - P.cellBase;
    const int32_t cx = (int32_t)(lc % P.gx), cy = (int32_t)((lc / P.gx) % P.gy), cz = (int32_t)(lc / (P.gx * P.gy));
    const int32_t xlo = maxi(cx - 1, 0), xhi = mini(cx + 1, (int32_t)P.gx - 1);
    const uint32_t p = plo + lane;
    const bool valid = p < phi;
    Dbg g;
    int3 pi = make_int3(0, 0, 0);
    uint32_t mi = 0;
    if (valid) {
      const int4 pr4 = D.prA[p];
      pi = make_int3(pr4.x, pr4.y, pr4.z);
      mi = D.matY[p];
    }
    DensAcc A{wmul(lutW[0], wadd(256, rhoW[mi])) >> 8, 0, 0, 0, 0};
    uint32_t cnt = 0, dropped = 0;
    int32_t bx0 = wave_min3(valid ? pi.x : I32_MAX) - (H - 1), bx1 = wave_max3(valid ? pi.x : I32_MIN) + (H - 1);
    int32_t by0 = wave_min3(valid ? pi.y : I32_MAX) - (H - 1), by1 = wave_max3(valid ? pi.y : I32_MIN) + (H - 1);
    int32_t bz0 = wave_min3(valid ? pi.z : I32_MAX) - (H - 1), bz1 = wave_max3(valid ? pi.z : I32_MIN) + (H - 1);
    uint64_t boxLanes = __ballot(valid);   // the lanes the reach box is taken over
    for (int32_t dz = -1; dz <= 1; ++dz) {
      const int32_t gz = cz + dz;
      if (gz < 0 || gz >= (int32_t)P.gz) continue;
      for (int32_t dy = -1; dy <= 1; ++dy) {
        const int32_t gy = cy + dy;
        if (gy < 0 || gy >= (int32_t)P.gy) continue;
        const uint32_t rowc = P.cellBase + ((uint32_t)gz * P.gy + (uint32_t)gy) * P.gx;
        const uint32_t lo = (uint32_t)((int32_t)D.cellStart[rowc + (uint32_t)xlo] + sh);
        const uint32_t hi = (uint32_t)((int32_t)D.cellStart[rowc + (uint32_t)xhi + 1u] + sh);
        for (uint32_t q0 = lo; q0 < hi; q0 += 64u) {
          if (!D.dbgOn) {
            // Without debug counters a full list has nothing left to find:
            // a wave whose lists are all full is done, and the reach box
            // shrinks to the lanes still filling.  In a packed cell the
            // lanes far from the pile fill last, and the pile's candidates
            // then drop out of the ballot instead of costing an iteration each.
            const bool open = valid && cnt < MAXNBR;
            const uint64_t openLanes = __ballot(open);
            if (openLanes == 0ull) break;
            if (openLanes != boxLanes) {
              boxLanes = openLanes;
              bx0 = wave_min3(open ? pi.x : I32_MAX) - (H - 1);
              bx1 = wave_max3(open ? pi.x : I32_MIN) + (H - 1);
              by0 = wave_min3(open ? pi.y : I32_MAX) - (H - 1);
              by1 = wave_max3(open ? pi.y : I32_MIN) + (H - 1);
              bz0 = wave_min3(open ? pi.z : I32_MAX) - (H - 1);
              bz1 = wave_max3(open ? pi.z : I32_MIN) + (H - 1);
            }
          }
          const uint32_t qn = (uint32_t)mini((int32_t)(hi - q0), 64);
          int4 cp = make_int4(0, 0, 0, 0);
          uint32_t cm = 0;
          if (lane < qn) {
            cp = D.prA[q0 + lane];
            cm = D.matY[q0 + lane];
          }
          // wait for this block's candidates here, so that the candidate
          // loop below never waits on the neighbour-list stores it issues
          __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
          const bool want = lane < qn && cp.x >= bx0 && cp.x <= bx1 && cp.y >= by0 && cp.y <= by1 &&
                            cp.z >= bz0 && cp.z <= bz1;
          const uint64_t todo = __ballot(want);
          if (todo == 0ull) continue;
          // the candidates in reach, in order, staged in the wave's LDS row as
          // {x, y, z, material | block lane << 8} and read back by every lane
          if (want)
            stg[__builtin_amdgcn_mbcnt_hi((uint32_t)(todo >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)todo, 0u))] =
                make_int4(cp.x, cp.y, cp.z, (int32_t)(cm | (lane << 8)));
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          const uint32_t m = (uint32_t)__builtin_popcountll(todo);
          for (uint32_t j = 0; j < m; ++j) {
            const int4 c = stg[j];
            const int32_t px = c.x, py = c.y, pz = c.z;
            const uint32_t mj = (uint32_t)c.w & 0xffu;
            const uint32_t q = q0 + ((uint32_t)c.w >> 8);
            if (!valid || q == p) continue;
            // Every lane's particle lies in this cell and the candidate inside
            // the reach box, so |d| < 2H - 1 on each axis: the quantised squared
            // radius cannot overflow, and it is >= H2Q whenever some |d| >= H
            // ((|d| >> 8)^2 >= 512^2), so the radius test alone decides, as the
            // reference's per-axis test plus radius test do (no |d| can be
            // INT_MIN for in-world positions, so fp_abs counts nothing).
            const int3 d = make_int3(wsub(px, pi.x), wsub(py, pi.y), wsub(pz, pi.z));
            const int32_t r2q = __mul24(d.x >> 8, d.x >> 8) + __mul24(d.y >> 8, d.y >> 8) + __mul24(d.z >> 8, d.z >> 8);
            if (r2q >= H2Q) continue;
            if (cnt >= MAXNBR) {
              dropped++;
              continue;
            }
            D.nbr[(size_t)(cnt++) * D.NS + p] = q;
            const uint32_t ui = (uint32_t)r2q >> 7;
            A.rho = wadd(A.rho, wmul(lutW[ui], wadd(256, rhoW[mj])) >> 8);
            const int32_t gt = lutG[ui];
            const int32_t gx = fp_mul_s(gt, d.x, 16, g), gy2 = fp_mul_s(gt, d.y, 16, g), gz2 = fp_mul_s(gt, d.z, 16, g);
            A.sx = wadd(A.sx, gx);
            A.sy = wadd(A.sy, gy2);
            A.sz = wadd(A.sz, gz2);
            A.S = wadd(wadd(wadd(A.S, s_term(gx, g)), s_term(gy2, g)), s_term(gz2, g));
          }
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");   // reads done before the next block's stores
        }
      }
    }
    if (valid) {
      D.nbrN[p] = cnt;
      D.drop[p] = D.dbgOn ? dropped : 0u;
      dens_finish(D, P, p, pi, A, D.prA, g);
    }
    flush_dbg3(D, s, g, 0);
  }
}

__global__ __launch_bounds__(256) void k3_density(Dev3 D, int4* pr) {
  uint32_t p = blockIdx.x * 256u + threadIdx.x;
  uint32_t s = p < D.N ? D.sidOf[p] : INACTIVE;
  Dbg g;
  if (s != INACTIVE) {
    const Scene3P& P = D.sp[s];
    int4 pr4 = pr[p];
    int3 pi = make_int3(pr4.x, pr4.y, pr4.z);
    DensAcc A{wmul(D.luts[OFF_W], wadd(256, mat_rhoW(D, D.matY[p]))) >> 8, 0, 0, 0, 0};
    uint32_t cnt = D.nbrN[p];
    const uint32_t* nb = D.nbr + p;
    for (uint32_t k = 0; k < cnt; ++k) {
      uint32_t q = nb[(size_t)k * D.NS];
      int4 pq = pr[q];
      int3 d = make_int3(wsub(pq.x, pi.x), wsub(pq.y, pi.y), wsub(pq.z, pi.z));
      int32_t r2q = r2q_of(d);
      if (r2q >= H2Q) continue;
      dens_term(D, A, d, (uint32_t)r2q >> 7, D.matY[q], g);
    }
    dens_finish(D, P, p, pi, A, pr, g);
  }
  flush_dbg3(D, s, g, 0);
}

// solveDp + applyDp (3d/sim3d.wgsl:476-537): reads `pr`, writes `out`
// (the other buffer).  On the last iteration: integrate and hash.
__global__ __launch_bounds__(256) void k3_solve_apply(Dev3 D, const int4* pr, int4* out, uint32_t last) {
  uint32_t p = blockIdx.x * 256u + threadIdx.x;
  uint32_t s = p < D.N ? D.sidOf[p] : INACTIVE;
  Dbg g;
  uint32_t hv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (s != INACTIVE) {
    const Scene3P& P = D.sp[s];
    int4 pr4 = pr[p];
    int3 pi = make_int3(pr4.x, pr4.y, pr4.z);
    int32_t lami = pr4.w;
    uint32_t cnt = D.nbrN[p];
    const uint32_t* nb = D.nbr + p;
    int32_t ax = 0, ay = 0, az = 0;
    for (uint32_t k = 0; k < cnt; ++k) {
      uint32_t q = nb[(size_t)k * D.NS];
      int4 pq = pr[q];
      int3 d = make_int3(wsub(pq.x, pi.x), wsub(pq.y, pi.y), wsub(pq.z, pi.z));
      int32_t r2q = r2q_of(d);
      if (r2q >= H2Q) continue;
      int32_t gt = D.luts[OFF_G + ((uint32_t)r2q >> 7)];
      int32_t gx = fp_mul_s(gt, d.x, 16, g), gy = fp_mul_s(gt, d.y, 16, g), gz = fp_mul_s(gt, d.z, 16, g);
      int32_t lsum = fp_add_sat(lami, pq.w, g);
      ax = fp_add_sat(ax, fp_mul_s(lsum, gx, 20, g), g);
      ay = fp_add_sat(ay, fp_mul_s(lsum, gy, 20, g), g);
      az = fp_add_sat(az, fp_mul_s(lsum, gz, 20, g), g);
    }
    for (uint32_t w = 0; w < 6; ++w) {
      int32_t dist = wall_dist(P, pi, w);
      if (dist >= H) continue;
      int3 gv = wall_dir(w, D.luts[OFF_PG + ((uint32_t)dist >> 8)], g);
      ax = fp_add_sat(ax, fp_mul_s(lami, gv.x, 20, g), g);
      ay = fp_add_sat(ay, fp_mul_s(lami, gv.y, 20, g), g);
      az = fp_add_sat(az, fp_mul_s(lami, gv.z, 20, g), g);
    }
    int32_t bw;
    int3 bg;
    if (ball_term(D, pi, bw, bg, g)) {
      ax = fp_add_sat(ax, fp_mul_s(lami, bg.x, 20, g), g);
      ay = fp_add_sat(ay, fp_mul_s(lami, bg.y, 20, g), g);
      az = fp_add_sat(az, fp_mul_s(lami, bg.z, 20, g), g);
    }
    int3 np = clamp_world(P, make_int3(fp_add_sat(pi.x, ax, g), fp_add_sat(pi.y, ay, g), fp_add_sat(pi.z, az, g)));
    out[p] = make_int4(np.x, np.y, np.z, lami);
    int4 ax4 = D.aux[p];
    D.aux[p] = make_int4(ax, ay, az, ax4.w);
    if (last) {
      int4 a = D.psY[p];
      int32_t vx = fp_sub_sat(np.x, a.x, g), vy = fp_sub_sat(np.y, a.y, g), vz = fp_sub_sat(np.z, a.z, g);
      uint32_t myid = D.idY[p];
      D.ps[p] = make_int4(np.x, np.y, np.z, a.w);
      D.vel[p] = make_int4(vx, vy, vz, 0);
      D.id[p] = myid;
      const uint32_t SEED[4] = {0x9e3779b1u, 0x85ebca77u, 0xc2b2ae3du, 0x27d4eb2fu};
      const uint32_t SALT[4] = {0x00000000u, 0x51633e2du, 0x9c8f2a17u, 0xd0e5b943u};
      uint32_t words[8] = {(uint32_t)np.x, (uint32_t)np.y, (uint32_t)np.z, (uint32_t)a.w,
                           (uint32_t)vx,   (uint32_t)vy,   (uint32_t)vz,   0u};
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        uint32_t h = SEED[l];
#pragma unroll
        for (int w = 0; w < 8; ++w) h = mm_step(h, words[w] ^ SALT[l]);
        hv[l] = fmix32(h ^ 8u);
        hv[4 + l] = fmix32(mm_step(h, myid) ^ 9u);
      }
    }
  }
  flush_dbg3(D, s, g, 0);
  if (!last) return;
  // Block partial (8 hash words + truncation count): a block lying wholly in
  // one scene stores its sums in hpart[block] for k3_hash_finish; a block
  // straddling scenes adds per lane with atomics.  All wrapping u32 sums.
  uint32_t h9[HPART3];
#pragma unroll
  for (int l = 0; l < 8; ++l) h9[l] = hv[l];
  h9[8] = s != INACTIVE ? D.drop[p] : 0u;
  __shared__ uint32_t wpart[4][HPART3];
  __shared__ uint32_t wscene[4];
  const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t s0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)s);
  const bool wuni = __all(s == s0);
#pragma unroll
  for (int l = 0; l < HPART3; ++l) {
    const uint32_t v = wave_sum3(h9[l]);
    if (lane == 0u) wpart[wv][l] = v;
  }
  if (lane == 0u) wscene[wv] = wuni ? s0 : MIXED;
  __syncthreads();
  const uint32_t bs = wscene[0];
  const bool buni = bs != MIXED && wscene[1] == bs && wscene[2] == bs && wscene[3] == bs;
  if (buni) {
    if (bs != INACTIVE && threadIdx.x < (uint32_t)HPART3) {
      const uint32_t t = threadIdx.x;
      D.hpart[(size_t)blockIdx.x * HPART3 + t] = wpart[0][t] + wpart[1][t] + wpart[2][t] + wpart[3][t];
    }
  } else if (s != INACTIVE) {
#pragma unroll
    for (int l = 0; l < 8; ++l) atomicAdd(&D.sd[s].acc[l], h9[l]);
    if (h9[8]) atomicAdd(&D.sd[s].dbg[6], h9[8]);
  }
}

// One wave per scene: fold the block partials of the blocks lying wholly in
// the scene's slot range into the atomically summed remainder, then the
// chain step, the H ring append and the substep counter.
__global__ __launch_bounds__(64) void k3_hash_finish(Dev3 D, const uint8_t* active) {
  const uint32_t s = blockIdx.x;
  if (!active[s]) return;
  Scene3D& SD = D.sd[s];
  const Scene3P& P = D.sp[s];
  const uint32_t lane = threadIdx.x;
  const uint32_t b0 = (P.partBase + 255u) / 256u, b1 = (P.partBase + P.n) / 256u;
  uint32_t sum[HPART3];
#pragma unroll
  for (int l = 0; l < HPART3; ++l) sum[l] = 0u;
  for (uint32_t b = b0 + lane; b < b1; b += 64u) {
#pragma unroll
    for (int l = 0; l < HPART3; ++l) sum[l] += D.hpart[(size_t)b * HPART3 + l];
  }
#pragma unroll
  for (int l = 0; l < HPART3; ++l) sum[l] = wave_sum3(sum[l]);
  if (lane != 0u) return;
  uint32_t a[8];
  for (int l = 0; l < 8; ++l) a[l] = SD.acc[l] + sum[l];
  SD.dbg[6] += sum[8];
  for (int j = 0; j < 8; ++j) SD.chain[j] = mm_step(SD.chain[j], a[j] ^ rotl32(a[(j + 1) & 7], 7));
  uint32_t* r = D.hring + ((size_t)s * D.ringCap + (SD.streamPos % D.ringCap)) * 8;
  for (int l = 0; l < 8; ++l) { r[l] = a[l]; SD.acc[l] = 0u; }
  SD.streamPos++;
  SD.substep++;
}

__global__ __launch_bounds__(256) void k3_hash_state(Dev3 D, uint32_t s) {
  const Scene3P& P = D.sp[s];
  uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= P.n) return;
  uint32_t p = P.partBase + t;
  int4 a = D.ps[p], v = D.vel[p];
  uint32_t myid = D.id[p];
  const uint32_t SEED[4] = {0x9e3779b1u, 0x85ebca77u, 0xc2b2ae3du, 0x27d4eb2fu};
  const uint32_t SALT[4] = {0x00000000u, 0x51633e2du, 0x9c8f2a17u, 0xd0e5b943u};
  uint32_t words[8] = {(uint32_t)a.x, (uint32_t)a.y, (uint32_t)a.z, (uint32_t)a.w,
                       (uint32_t)v.x, (uint32_t)v.y, (uint32_t)v.z, (uint32_t)v.w};
  for (int l = 0; l < 4; ++l) {
    uint32_t h = SEED[l];
    for (int w = 0; w < 8; ++w) h = mm_step(h, words[w] ^ SALT[l]);
    atomicAdd(&D.hacc[s * 8 + l], fmix32(h ^ 8u));
    atomicAdd(&D.hacc[s * 8 + 4 + l], fmix32(mm_step(h, myid) ^ 9u));
  }
}

// Full 16-word records in slot order: p, v, pr (pred, lambda), dp (dp, rho).
__global__ __launch_bounds__(256) void k3_to_slots(Dev3 D, uint32_t s, int32_t* out16, const int4* prLast) {
  const Scene3P& P = D.sp[s];
  uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= P.n) return;
  uint32_t p = P.partBase + t;
  uint32_t i = D.id[p];
  int4 a = D.ps[p], v = D.vel[p];
  int32_t* o = out16 + (size_t)i * 16;
  o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
  o[4] = v.x; o[5] = v.y; o[6] = v.z; o[7] = v.w;
  if (prLast && D.sd[s].substep > 0) {
    int4 pr = prLast[p], ax = D.aux[p];
    o[8] = pr.x; o[9] = pr.y; o[10] = pr.z; o[11] = pr.w;
    o[12] = ax.x; o[13] = ax.y; o[14] = ax.z; o[15] = ax.w;
  } else {
    for (int k = 8; k < 16; ++k) o[k] = 0;
  }
}

// The reference's striped FNV over slot-ordered records (3d/sim3d.wgsl:571-590).
__global__ void k3_fnv_stripes(const int32_t* rec16, uint32_t n, uint32_t* words) {
  uint32_t k = threadIdx.x;
  uint32_t per = (n + 255u) / 256u;
  uint32_t lo = k * per, hi = min(lo + per, n);
  uint32_t h = 2166136261u;
  for (uint32_t i = lo; i < hi; ++i) {
    const int32_t* r = rec16 + (size_t)i * 16;
    const int idx[6] = {0, 1, 2, 4, 5, 6};
    for (int w = 0; w < 6; ++w) {
      uint32_t x = (uint32_t)r[idx[w]];
      h = (h ^ (x & 0xffu)) * 16777619u;
      h = (h ^ ((x >> 8) & 0xffu)) * 16777619u;
      h = (h ^ ((x >> 16) & 0xffu)) * 16777619u;
      h = (h ^ ((x >> 24) & 0xffu)) * 16777619u;
    }
  }
  words[k] = h;
}

// 3D stir (this build's definition; the reference has no 3D impulse): fluid
// slots in ascending order within r of (cx, cy, cz), IEEE double test, at
// most `limit`, get (ix, iy, iz) added to v, clamped to +-65536.
__global__ __launch_bounds__(256) void k3_impulse_mark(Dev3 D, uint32_t group) {
  uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p >= D.N) return;
  uint32_t s = D.sidOf[p];
  if (s == INACTIVE) return;
  const Scene3P& P = D.sp[s];
  const Scene3D& SD = D.sd[s];
  uint32_t id = D.id[p];
  uint32_t e = SD.evCursor + group, hit = 0;
  if (e < SD.evBase + SD.evCount && D.events[e].substep == SD.substep) {
    const Event3& E = D.events[e];
    int4 a = D.ps[p];
    double dx = (double)a.x - (double)E.cx, dy = (double)a.y - (double)E.cy, dz = (double)a.z - (double)E.cz;
    double q = dx * dx;
    q = q + dy * dy;
    q = q + dz * dz;
    double r2 = (double)E.r * (double)E.r;
    hit = q > r2 ? 0u : 1u;
  }
  D.hitFlag[P.partBase + id] = hit;
}
__global__ __launch_bounds__(256) void k3_impulse_apply(Dev3 D, uint32_t group) {
  uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p >= D.N) return;
  uint32_t s = D.sidOf[p];
  if (s == INACTIVE) return;
  const Scene3P& P = D.sp[s];
  const Scene3D& SD = D.sd[s];
  uint32_t id = D.id[p];
  uint32_t e = SD.evCursor + group;
  if (e >= SD.evBase + SD.evCount || D.events[e].substep != SD.substep) return;
  uint32_t slot = P.partBase + id;
  if (!D.hitFlag[slot]) return;
  const Event3& E = D.events[e];
  if (D.hitScan[slot] - D.hitScan[P.partBase] >= E.limit) return;
  int4 v = D.vel[p];
  auto cl = [](int64_t x) { return (int32_t)(x < -65536 ? -65536 : (x > 65536 ? 65536 : x)); };
  v.x = cl((int64_t)v.x + E.ix);
  v.y = cl((int64_t)v.y + E.iy);
  v.z = cl((int64_t)v.z + E.iz);
  D.vel[p] = v;
}
__global__ void k3_impulse_advance(Dev3 D, const uint8_t* active, uint32_t groups) {
  uint32_t s = blockIdx.x * 64u + threadIdx.x;
  if (s >= D.S || !active[s]) return;
  Scene3D& SD = D.sd[s];
  uint32_t end = SD.evBase + SD.evCount, k = 0;
  while (k < groups && SD.evCursor < end && D.events[SD.evCursor].substep == SD.substep) { SD.evCursor++; k++; }
  while (SD.evCursor < end && D.events[SD.evCursor].substep < SD.substep) SD.evCursor++;
}

inline uint32_t div_up3(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

}  // namespace

struct sn3_engine {
  sn_opts opts{};
  hipStream_t stream = nullptr;
  std::string err;
  Dev3 D{};
  std::vector<Scene3P> sp;
  std::vector<uint8_t> active;
  uint8_t* dActive = nullptr;
  uint32_t S = 0, N = 0, C = 0;
  std::vector<void*> allocs;
  hipGraphExec_t graph = nullptr;
  uint32_t graphSubsteps = 0;
  bool loaded = false;
  bool slotted = false;
  uint32_t nbrBlocks = 8;    // persistent grid of the neighbour sweep
  uint32_t maxEvents = 0;
  Event3* dEvents = nullptr;
  uint32_t eventCap = 0;
  hipEvent_t ev[6];
};

namespace {

template <class T>
T* dalloc3(sn3_engine* e, size_t count) {
  void* p = nullptr;
  HIPCHECK3(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
  e->allocs.push_back(p);
  return static_cast<T*>(p);
}
int fail3(sn3_engine* e, int code, const std::string& m) {
  if (e) e->err = m;
  return code;
}
void drop_graph3(sn3_engine* e) {
  if (e->graph) {
    hipGraphExecDestroy(e->graph);
    e->graph = nullptr;
  }
}
void scan3(sn3_engine* e, uint32_t* count, uint32_t n, uint32_t* out, bool clear) {
  uint32_t nb = div_up3(std::max(n, 1u), SCAN_BLOCK);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, e->stream, count, n, e->D.blockSums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, e->stream, e->D.blockSums, nb);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(256), 0, e->stream, count, n, e->D.blockSums, out, clear);
}
void enqueue_substep3(sn3_engine* e, hipEvent_t* marks) {
  Dev3& D = e->D;
  hipStream_t st = e->stream;
  const uint32_t gN = div_up3(std::max(D.N, 1u), 256);
  for (uint32_t gi = 0; gi < e->maxEvents; ++gi) {
    hipLaunchKernelGGL(k3_impulse_mark, dim3(gN), dim3(256), 0, st, D, gi);
    scan3(e, D.hitFlag, D.N, D.hitScan, false);
    hipLaunchKernelGGL(k3_impulse_apply, dim3(gN), dim3(256), 0, st, D, gi);
  }
  if (e->maxEvents)
    hipLaunchKernelGGL(k3_impulse_advance, dim3(div_up3(e->S, 64)), dim3(64), 0, st, D, e->dActive, e->maxEvents);
  if (marks) hipEventRecord(marks[0], st);
  hipLaunchKernelGGL(k3_predict_count, dim3(gN), dim3(256), 0, st, D);
  scan3(e, D.cellCount, D.C, D.cellStart, true);
  hipLaunchKernelGGL(k3_scatter, dim3(gN), dim3(256), 0, st, D);
  hipLaunchKernelGGL(k3_canon_gather, dim3(gN), dim3(256), 0, st, D);
  if (marks) hipEventRecord(marks[1], st);
  hipLaunchKernelGGL(k3_build_lists, dim3(div_up3(std::max(D.C, 1u), 256)), dim3(256), 0, st, D);
  hipLaunchKernelGGL(k3_nbr_tile, dim3(e->nbrBlocks), dim3(256), 0, st, D);
  if (marks) hipEventRecord(marks[2], st);
  int4* cur = D.prA;
  int4* nxt = D.prB;
  const uint32_t iters = e->opts.iters;
  for (uint32_t it = 0; it < iters; ++it) {
    if (it > 0) hipLaunchKernelGGL(k3_density, dim3(gN), dim3(256), 0, st, D, cur);
    hipLaunchKernelGGL(k3_solve_apply, dim3(gN), dim3(256), 0, st, D, (const int4*)cur, nxt,
                       (uint32_t)(it + 1 == iters));
    std::swap(cur, nxt);
  }
  if (marks) hipEventRecord(marks[3], st);
  hipLaunchKernelGGL(k3_hash_finish, dim3(e->S), dim3(64), 0, st, D, e->dActive);
  if (marks) hipEventRecord(marks[4], st);
}
int4* last_pred3(sn3_engine* e) { return (e->opts.iters % 2 == 1) ? e->D.prB : e->D.prA; }

void upload_layout3(sn3_engine* e) {
  std::vector<uint32_t> sid(e->N, INACTIVE);
  uint32_t before = 0;
  for (uint32_t s = 0; s < e->S; ++s) {
    Scene3P& P = e->sp[s];
    P.posShift = (int32_t)P.partBase - (int32_t)before;
    if (e->active[s]) {
      for (uint32_t k = 0; k < P.n; ++k) sid[P.partBase + k] = s;
      before += P.n;
    }
  }
  HIPCHECK3(hipMemcpy((void*)e->D.sidOf, sid.data(), 4 * (size_t)e->N, hipMemcpyHostToDevice));
  std::vector<uint32_t> cs(std::max(e->C, 1u), INACTIVE);
  for (uint32_t s = 0; s < e->S; ++s) {
    if (!e->active[s]) continue;
    const Scene3P& P = e->sp[s];
    for (uint32_t k = 0; k < P.cellTotal; ++k) cs[P.cellBase + k] = s;
  }
  HIPCHECK3(hipMemcpy((void*)e->D.cellScene, cs.data(), 4 * (size_t)e->C, hipMemcpyHostToDevice));
  HIPCHECK3(hipMemcpy((void*)e->D.sp, e->sp.data(), sizeof(Scene3P) * e->S, hipMemcpyHostToDevice));
  HIPCHECK3(hipMemcpy(e->dActive, e->active.data(), e->S, hipMemcpyHostToDevice));
}

int read_sd3(sn3_engine* e, uint32_t scene, Scene3D& out) {
  if (!e->loaded) return fail3(e, SN_E_STATE, "[aether-3d] no scene loaded");
  if (scene >= e->S) return fail3(e, SN_E_ARG, "[aether-3d] scene index out of range");
  HIPCHECK3(hipSetDevice((int)e->opts.device));
  HIPCHECK3(hipStreamSynchronize(e->stream));
  HIPCHECK3(hipMemcpy(&out, e->D.sd + scene, sizeof(Scene3D), hipMemcpyDeviceToHost));
  return SN_OK;
}

int read_records3(sn3_engine* e, uint32_t scene, int32_t* out16) {
  const Scene3P& P = e->sp[scene];
  int32_t* d = nullptr;
  HIPCHECK3(hipMalloc((void**)&d, std::max<size_t>(P.n, 1) * 64));
  hipLaunchKernelGGL(k3_to_slots, dim3(div_up3(std::max(P.n, 1u), 256)), dim3(256), 0, e->stream, e->D, scene, d,
                     (const int4*)last_pred3(e));
  HIPCHECK3(hipMemcpyAsync(out16, d, (size_t)P.n * 64, hipMemcpyDeviceToHost, e->stream));
  HIPCHECK3(hipStreamSynchronize(e->stream));
  hipFree(d);
  return SN_OK;
}

}  // namespace

extern "C" {

int sn3_create(const sn_opts* opts, const int32_t* luts3, uint32_t lut_words, sn3_engine** out) {
  if (!opts || !luts3 || !out) return SN_E_ARG;
  sn3_engine* e = new sn3_engine();
  *out = e;
  e->opts = *opts;
  if (!e->opts.max_scenes) e->opts.max_scenes = 1;
  if (!e->opts.iters) e->opts.iters = 3;
  if (!e->opts.stream_cap) e->opts.stream_cap = 4096;
  if (lut_words < LUT3_WORDS) return fail3(e, SN_E_ARG, "[aether-3d] table image too short");
  if (e->opts.iters > 16) return fail3(e, SN_E_ARG, "[aether-3d] iters must be in 1..16");
  try {
    HIPCHECK3(hipSetDevice((int)e->opts.device));
    HIPCHECK3(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    for (auto& ev : e->ev) HIPCHECK3(hipEventCreate(&ev));
    const uint32_t NC = e->opts.max_particles, CC = e->opts.max_cells, SC = e->opts.max_scenes;
    if (std::max(NC, CC) > 4096u * SCAN_BLOCK) return fail3(e, SN_E_CAPACITY, "[aether-3d] capacity above 16M");
    Dev3& D = e->D;
    D.dbgOn = e->opts.no_debug ? 0u : 1u;
    D.ps = dalloc3<int4>(e, NC);
    D.vel = dalloc3<int4>(e, NC);
    D.id = dalloc3<uint32_t>(e, NC);
    D.psY = dalloc3<int4>(e, NC);
    D.idY = dalloc3<uint32_t>(e, NC);
    D.matY = dalloc3<uint32_t>(e, NC);
    D.sidOf = dalloc3<uint32_t>(e, NC);
    D.pr0 = dalloc3<int4>(e, NC);
    D.cell0 = dalloc3<uint32_t>(e, NC);
    D.rank0 = dalloc3<uint32_t>(e, NC);
    D.bucket = dalloc3<uint32_t>(e, NC);
    D.bucketId = dalloc3<uint32_t>(e, NC);
    D.cellCount = dalloc3<uint32_t>(e, CC);
    D.cellStart = dalloc3<uint32_t>(e, (size_t)CC + 1);
    D.blockSums = dalloc3<uint32_t>(e, div_up3(std::max(CC, NC), SCAN_BLOCK) + 1);
    D.cellY = dalloc3<uint32_t>(e, NC);
    D.prA = dalloc3<int4>(e, NC);
    D.prB = dalloc3<int4>(e, NC);
    D.aux = dalloc3<int4>(e, NC);
    D.nbr = dalloc3<uint32_t>(e, (size_t)NC * MAXNBR);
    D.NS = NC;
    D.nbrN = dalloc3<uint32_t>(e, NC);
    D.drop = dalloc3<uint32_t>(e, NC);
    D.hpart = dalloc3<uint32_t>(e, (size_t)div_up3(NC, 256) * HPART3);
    D.cellScene = dalloc3<uint32_t>(e, CC);
    // at most one item per cell plus one per full chunk
    D.dlist = dalloc3<uint2>(e, (size_t)CC + NC / CHUNK3 + 64);
    D.listCount = dalloc3<uint32_t>(e, 8);
    {
      hipDeviceProp_t prop;
      HIPCHECK3(hipGetDeviceProperties(&prop, (int)e->opts.device));
      int bN = 0;
      HIPCHECK3(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bN, reinterpret_cast<const void*>(k3_nbr_tile), 256, 0));
      const uint32_t cus = (uint32_t)std::max(prop.multiProcessorCount, 1);
      e->nbrBlocks = std::max(8u, (cus * (uint32_t)std::max(bN, 1)) & ~7u);
    }
    D.hitFlag = dalloc3<uint32_t>(e, NC);
    D.hitScan = dalloc3<uint32_t>(e, (size_t)NC + 1);
    int32_t* l = dalloc3<int32_t>(e, LUT3_WORDS);
    HIPCHECK3(hipMemcpy(l, luts3, 4 * LUT3_WORDS, hipMemcpyHostToDevice));
    D.luts = l;
    D.sp = dalloc3<Scene3P>(e, SC);
    D.sd = dalloc3<Scene3D>(e, SC);
    D.hacc = dalloc3<uint32_t>(e, (size_t)SC * 8);
    D.hring = dalloc3<uint32_t>(e, (size_t)SC * e->opts.stream_cap * 8);
    D.ringCap = e->opts.stream_cap;
    e->dActive = dalloc3<uint8_t>(e, SC);
    HIPCHECK3(hipMemset(D.cellCount, 0, 4 * (size_t)CC));
    std::memset(&D.in, 0, sizeof(D.in));
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

void sn3_destroy(sn3_engine* e) {
  if (!e) return;
  hipSetDevice((int)e->opts.device);
  if (e->stream) hipStreamSynchronize(e->stream);
  drop_graph3(e);
  for (void* p : e->allocs) hipFree(p);
  if (e->dEvents) hipFree(e->dEvents);
  for (auto& ev : e->ev)
    if (ev) hipEventDestroy(ev);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
}

const char* sn3_last_error(const sn3_engine* e) { return e ? e->err.c_str() : "null engine"; }

int sn3_load(sn3_engine* e, const sn3_scene* scenes, uint32_t count) {
  if (!e || !scenes || !count) return SN_E_ARG;
  if (count > e->opts.max_scenes) return fail3(e, SN_E_CAPACITY, "[aether-3d] too many scenes");
  uint64_t totN = 0, totC = 0;
  std::vector<Scene3P> sp(count);
  for (uint32_t s = 0; s < count; ++s) {
    const sn3_scene& sc = scenes[s];
    if (!sc.gx || !sc.gy || !sc.gz) return fail3(e, SN_E_ARG, "[aether-3d] empty grid");
    if ((uint64_t)sc.gx * H > 0x7fffffffull || (uint64_t)sc.gy * H > 0x7fffffffull ||
        (uint64_t)sc.gz * H > 0x7fffffffull)
      return fail3(e, SN_E_ARG, "[aether-3d] world extent exceeds i32 ticks");
    for (uint32_t i = 0; i < sc.n; ++i) {
      const int32_t* r = sc.parts16 + (size_t)i * 16;
      if ((uint32_t)r[3] >= NMAT3) return fail3(e, SN_E_ARG, "[aether-3d] material outside the table");
    }
    Scene3P& P = sp[s];
    std::memset(&P, 0, sizeof(P));
    P.n = sc.n;
    P.gx = sc.gx;
    P.gy = sc.gy;
    P.gz = sc.gz;
    P.cellTotal = sc.gx * sc.gy * sc.gz;
    P.cellBase = (uint32_t)totC;
    P.partBase = (uint32_t)totN;
    P.wx = (int32_t)sc.gx * H;
    P.wy = (int32_t)sc.gy * H;
    P.wz = (int32_t)sc.gz * H;
    totN += sc.n;
    totC += P.cellTotal;
  }
  if (totN > e->opts.max_particles) return fail3(e, SN_E_CAPACITY, "[aether-3d] particles exceed maxParticles");
  if (totC > e->opts.max_cells) return fail3(e, SN_E_CAPACITY, "[aether-3d] cells exceed maxCells");
  try {
    HIPCHECK3(hipSetDevice((int)e->opts.device));
    HIPCHECK3(hipStreamSynchronize(e->stream));
    drop_graph3(e);
    e->S = count;
    e->N = (uint32_t)totN;
    e->C = (uint32_t)totC;
    e->sp = sp;
    e->active.assign(count, 1);
    e->slotted = false;
    Dev3& D = e->D;
    D.N = e->N;
    D.C = e->C;
    D.S = e->S;
    std::vector<int4> ps(e->N), vel(e->N);
    std::vector<uint32_t> id(e->N);
    for (uint32_t s = 0; s < count; ++s)
      for (uint32_t i = 0; i < scenes[s].n; ++i) {
        const int32_t* r = scenes[s].parts16 + (size_t)i * 16;
        uint32_t p = sp[s].partBase + i;
        ps[p] = make_int4(r[0], r[1], r[2], r[3]);
        vel[p] = make_int4(r[4], r[5], r[6], r[7]);
        id[p] = i;
      }
    if (e->N) {
      HIPCHECK3(hipMemcpy(D.ps, ps.data(), 16 * (size_t)e->N, hipMemcpyHostToDevice));
      HIPCHECK3(hipMemcpy(D.vel, vel.data(), 16 * (size_t)e->N, hipMemcpyHostToDevice));
      HIPCHECK3(hipMemcpy(D.id, id.data(), 4 * (size_t)e->N, hipMemcpyHostToDevice));
    }
    std::vector<Scene3D> sd(count);
    static const uint32_t CHAIN_INIT[8] = {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u,
                                           0xa4093822u, 0x299f31d0u, 0x082efa98u, 0xec4e6c89u};
    for (auto& d : sd) {
      std::memset(&d, 0, sizeof(d));
      for (int j = 0; j < 8; ++j) d.chain[j] = CHAIN_INIT[j];
    }
    HIPCHECK3(hipMemcpy(D.sd, sd.data(), sizeof(Scene3D) * count, hipMemcpyHostToDevice));
    HIPCHECK3(hipMemset(D.cellCount, 0, 4 * (size_t)std::max(e->C, 1u)));
    e->maxEvents = 0;
    upload_layout3(e);
    e->loaded = true;
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

// Slotted 3D worlds (continuous batching): see sn_reserve_slots.
int sn3_reserve_slots(sn3_engine* e, uint32_t count, const uint32_t* cap_particles, const uint32_t* cap_cells) {
  if (!e || !cap_particles || !cap_cells || !count) return SN_E_ARG;
  if (count > e->opts.max_scenes) return fail3(e, SN_E_CAPACITY, "[aether-3d] too many slots");
  uint64_t totN = 0, totC = 0;
  std::vector<Scene3P> sp(count);
  for (uint32_t s = 0; s < count; ++s) {
    Scene3P& P = sp[s];
    std::memset(&P, 0, sizeof(P));
    P.partBase = (uint32_t)totN;
    P.cellBase = (uint32_t)totC;
    P.cap = cap_particles[s];
    P.cellCap = cap_cells[s];
    P.gx = P.gy = P.gz = 1;
    P.wx = P.wy = P.wz = H;
    totN += cap_particles[s];
    totC += cap_cells[s];
  }
  if (totN > e->opts.max_particles || totC > e->opts.max_cells)
    return fail3(e, SN_E_CAPACITY, "[aether-3d] slot capacities exceed maxParticles / maxCells");
  try {
    HIPCHECK3(hipSetDevice((int)e->opts.device));
    HIPCHECK3(hipStreamSynchronize(e->stream));
    drop_graph3(e);
    e->S = count;
    e->N = (uint32_t)totN;
    e->C = (uint32_t)totC;
    e->sp = sp;
    e->active.assign(count, 0);
    Dev3& D = e->D;
    D.N = e->N;
    D.C = e->C;
    D.S = e->S;
    std::vector<Scene3D> sd(count);
    for (auto& d : sd) std::memset(&d, 0, sizeof(d));
    HIPCHECK3(hipMemcpy(D.sd, sd.data(), sizeof(Scene3D) * count, hipMemcpyHostToDevice));
    HIPCHECK3(hipMemset(D.cellCount, 0, 4 * (size_t)std::max(e->C, 1u)));
    e->maxEvents = 0;
    upload_layout3(e);
    e->loaded = true;
    e->slotted = true;
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn3_load_slot(sn3_engine* e, uint32_t slot, const sn3_scene* sc) {
  if (!e || !sc) return SN_E_ARG;
  if (!e->loaded || !e->slotted) return fail3(e, SN_E_STATE, "[aether-3d] sn3_load_slot needs a slotted world");
  if (slot >= e->S) return fail3(e, SN_E_ARG, "[aether-3d] slot index out of range");
  if (!sc->gx || !sc->gy || !sc->gz) return fail3(e, SN_E_ARG, "[aether-3d] empty grid");
  if ((uint64_t)sc->gx * H > 0x7fffffffull || (uint64_t)sc->gy * H > 0x7fffffffull ||
      (uint64_t)sc->gz * H > 0x7fffffffull)
    return fail3(e, SN_E_ARG, "[aether-3d] world extent exceeds i32 ticks");
  Scene3P P = e->sp[slot];
  const uint64_t cells = (uint64_t)sc->gx * sc->gy * sc->gz;
  if (sc->n > P.cap || cells > P.cellCap)
    return fail3(e, SN_E_CAPACITY, "[aether-3d] scene does not fit slot " + std::to_string(slot));
  for (uint32_t i = 0; i < sc->n; ++i)
    if ((uint32_t)sc->parts16[(size_t)i * 16 + 3] >= NMAT3)
      return fail3(e, SN_E_ARG, "[aether-3d] material outside the table");
  try {
    HIPCHECK3(hipSetDevice((int)e->opts.device));
    HIPCHECK3(hipStreamSynchronize(e->stream));
    drop_graph3(e);
    P.n = sc->n;
    P.gx = sc->gx;
    P.gy = sc->gy;
    P.gz = sc->gz;
    P.cellTotal = (uint32_t)cells;
    P.wx = (int32_t)sc->gx * H;
    P.wy = (int32_t)sc->gy * H;
    P.wz = (int32_t)sc->gz * H;
    e->sp[slot] = P;
    Dev3& D = e->D;
    std::vector<int4> ps(sc->n), vel(sc->n);
    std::vector<uint32_t> id(sc->n);
    for (uint32_t i = 0; i < sc->n; ++i) {
      const int32_t* r = sc->parts16 + (size_t)i * 16;
      ps[i] = make_int4(r[0], r[1], r[2], r[3]);
      vel[i] = make_int4(r[4], r[5], r[6], r[7]);
      id[i] = i;
    }
    if (sc->n) {
      HIPCHECK3(hipMemcpy(D.ps + P.partBase, ps.data(), 16 * (size_t)sc->n, hipMemcpyHostToDevice));
      HIPCHECK3(hipMemcpy(D.vel + P.partBase, vel.data(), 16 * (size_t)sc->n, hipMemcpyHostToDevice));
      HIPCHECK3(hipMemcpy(D.id + P.partBase, id.data(), 4 * (size_t)sc->n, hipMemcpyHostToDevice));
    }
    static const uint32_t CHAIN_INIT[8] = {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u,
                                           0xa4093822u, 0x299f31d0u, 0x082efa98u, 0xec4e6c89u};
    Scene3D d;
    std::memset(&d, 0, sizeof(d));
    for (int j = 0; j < 8; ++j) d.chain[j] = CHAIN_INIT[j];
    HIPCHECK3(hipMemcpy(D.sd + slot, &d, sizeof(Scene3D), hipMemcpyHostToDevice));
    e->active[slot] = 1;
    upload_layout3(e);
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn3_set_inputs(sn3_engine* e, int32_t stir_amp, int32_t stir_phase, uint32_t ball_on, const int32_t* ball3,
                   int32_t ball_reach) {
  if (!e) return SN_E_ARG;
  e->D.in.stirAmp = stir_amp;
  e->D.in.stirPhase = stir_phase;
  e->D.in.ballOn = ball_on;
  if (ball3) {
    e->D.in.ballX = ball3[0];
    e->D.in.ballY = ball3[1];
    e->D.in.ballZ = ball3[2];
  }
  e->D.in.ballReach = ball_reach;
  drop_graph3(e);
  return SN_OK;
}

int sn3_step(sn3_engine* e, uint32_t count) {
  if (!e) return SN_E_ARG;
  if (!e->loaded) return fail3(e, SN_E_STATE, "[aether-3d] step before load");
  try {
    HIPCHECK3(hipSetDevice((int)e->opts.device));
    const uint32_t G = 8;
    if (!e->graph) {
      hipGraph_t g;
      HIPCHECK3(hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
      try {
        for (uint32_t k = 0; k < G; ++k) enqueue_substep3(e, nullptr);
      } catch (const DevError3&) {
        hipGraph_t broken = nullptr;
        (void)hipStreamEndCapture(e->stream, &broken);   // leave the stream out of capture
        if (broken) (void)hipGraphDestroy(broken);
        throw;
      }
      HIPCHECK3(hipStreamEndCapture(e->stream, &g));
      HIPCHECK3(hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0));
      HIPCHECK3(hipGraphDestroy(g));
      e->graphSubsteps = G;
    }
    uint32_t done = 0;
    while (count - done >= e->graphSubsteps) {
      HIPCHECK3(hipGraphLaunch(e->graph, e->stream));
      done += e->graphSubsteps;
    }
    for (; done < count; ++done) enqueue_substep3(e, nullptr);
    HIPCHECK3(hipGetLastError());
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn3_query(sn3_engine* e) {
  if (!e) return SN_E_ARG;
  hipError_t q = hipStreamQuery(e->stream);
  if (q == hipSuccess) return SN_OK;
  if (q == hipErrorNotReady) return 1;
  return fail3(e, SN_E_DEVICE, std::string("hipStreamQuery: ") + hipGetErrorString(q));
}

int sn3_sync(sn3_engine* e) {
  if (!e) return SN_E_ARG;
  try {
    HIPCHECK3(hipSetDevice((int)e->opts.device));
    HIPCHECK3(hipStreamSynchronize(e->stream));
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn3_read_hash(sn3_engine* e, uint32_t scene, uint32_t acc_out[8]) {
  if (!e || !acc_out) return SN_E_ARG;
  try {
    Scene3D sd;
    int rc = read_sd3(e, scene, sd);
    if (rc) return rc;
    HIPCHECK3(hipMemsetAsync(e->D.hacc, 0, 32 * (size_t)e->S, e->stream));
    hipLaunchKernelGGL(k3_hash_state, dim3(div_up3(std::max(e->sp[scene].n, 1u), 256)), dim3(256), 0, e->stream, e->D,
                       scene);
    HIPCHECK3(hipMemcpyAsync(acc_out, e->D.hacc + 8 * scene, 32, hipMemcpyDeviceToHost, e->stream));
    HIPCHECK3(hipStreamSynchronize(e->stream));
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn3_read_stream(sn3_engine* e, uint32_t scene, uint32_t* out, size_t max_substeps, size_t* n_out) {
  if (!e || !n_out || (!out && max_substeps)) return SN_E_ARG;
  try {
    Scene3D sd;
    int rc = read_sd3(e, scene, sd);
    if (rc) return rc;
    if (sd.streamPos > e->D.ringCap) return fail3(e, SN_E_CAPACITY, "[aether-3d] H ring overflowed");
    size_t n = std::min<size_t>(sd.streamPos, max_substeps);
    if (n) HIPCHECK3(hipMemcpy(out, e->D.hring + (size_t)scene * e->D.ringCap * 8, n * 32, hipMemcpyDeviceToHost));
    *n_out = n;
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn3_reset_stream(sn3_engine* e) {
  if (!e || !e->loaded) return SN_E_ARG;
  try {
    HIPCHECK3(hipStreamSynchronize(e->stream));
    std::vector<Scene3D> sd(e->S);
    HIPCHECK3(hipMemcpy(sd.data(), e->D.sd, sizeof(Scene3D) * e->S, hipMemcpyDeviceToHost));
    for (auto& d : sd) d.streamPos = 0;
    HIPCHECK3(hipMemcpy(e->D.sd, sd.data(), sizeof(Scene3D) * e->S, hipMemcpyHostToDevice));
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn3_read_chain(sn3_engine* e, uint32_t scene, uint32_t chain_out[8]) {
  if (!e || !chain_out) return SN_E_ARG;
  try {
    Scene3D sd;
    int rc = read_sd3(e, scene, sd);
    if (rc) return rc;
    for (int j = 0; j < 8; ++j) chain_out[j] = sd.chain[j];
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn3_debug_counters(sn3_engine* e, uint32_t scene, uint32_t out[16]) {
  if (!e || !out) return SN_E_ARG;
  try {
    Scene3D sd;
    int rc = read_sd3(e, scene, sd);
    if (rc) return rc;
    for (int j = 0; j < 16; ++j) out[j] = sd.dbg[j];
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn3_read_state(sn3_engine* e, uint32_t scene, int32_t* out16, size_t words) {
  if (!e || !out16) return SN_E_ARG;
  try {
    Scene3D sd;
    int rc = read_sd3(e, scene, sd);
    if (rc) return rc;
    if (words < (size_t)e->sp[scene].n * 16) return fail3(e, SN_E_ARG, "[aether-3d] buffer too small");
    return read_records3(e, scene, out16);
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
}

int sn3_fnv_digest(sn3_engine* e, uint32_t scene, uint32_t lanes_out[8]) {
  if (!e || !lanes_out) return SN_E_ARG;
  try {
    Scene3D sd;
    int rc = read_sd3(e, scene, sd);
    if (rc) return rc;
    const Scene3P& P = e->sp[scene];
    int32_t* rec = nullptr;
    uint32_t* words = nullptr;
    HIPCHECK3(hipMalloc((void**)&rec, std::max<size_t>(P.n, 1) * 64));
    HIPCHECK3(hipMalloc((void**)&words, 256 * 4));
    hipLaunchKernelGGL(k3_to_slots, dim3(div_up3(std::max(P.n, 1u), 256)), dim3(256), 0, e->stream, e->D, scene, rec,
                       (const int4*)nullptr);
    hipLaunchKernelGGL(k3_fnv_stripes, dim3(1), dim3(256), 0, e->stream, (const int32_t*)rec, P.n, words);
    uint32_t w[256];
    HIPCHECK3(hipMemcpyAsync(w, words, 1024, hipMemcpyDeviceToHost, e->stream));
    HIPCHECK3(hipStreamSynchronize(e->stream));
    hipFree(rec);
    hipFree(words);
    // host fold in index order (3d/app.js:731-742)
    uint32_t lanes[8];
    for (uint32_t L = 0; L < 8; ++L) lanes[L] = 2166136261u ^ (uint32_t)((uint64_t)L * 0x9e3779b9ull);
    for (uint32_t k = 0; k < 256; ++k)
      for (uint32_t L = 0; L < 8; ++L) {
        uint32_t h = (lanes[L] ^ w[k]) * 16777619u;
        lanes[L] = h ^ (h >> 13);
      }
    for (int L = 0; L < 8; ++L) lanes_out[L] = lanes[L];
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn3_set_events(sn3_engine* e, const sn3_event* events, uint32_t count) {
  if (!e || (!events && count)) return SN_E_ARG;
  if (!e->loaded) return fail3(e, SN_E_STATE, "[aether-3d] no scene loaded");
  try {
    HIPCHECK3(hipStreamSynchronize(e->stream));
    std::vector<Scene3D> sd(e->S);
    HIPCHECK3(hipMemcpy(sd.data(), e->D.sd, sizeof(Scene3D) * e->S, hipMemcpyDeviceToHost));
    std::vector<Event3> ev(count);
    uint32_t maxGroup = 0;
    for (uint32_t k = 0; k < count; ++k) {
      const sn3_event& x = events[k];
      if (x.scene >= e->S) return fail3(e, SN_E_ARG, "[aether-3d] event scene out of range");
      if (k && (x.scene < events[k - 1].scene ||
                (x.scene == events[k - 1].scene && x.substep < events[k - 1].substep)))
        return fail3(e, SN_E_ARG, "[aether-3d] events must be sorted by (scene, substep)");
      ev[k] = Event3{x.scene, x.substep, x.cx, x.cy, x.cz, x.r, x.ix, x.iy, x.iz, x.limit};
    }
    for (auto& d : sd) d.evBase = d.evCount = d.evCursor = 0;
    for (uint32_t k = 0; k < count;) {
      uint32_t s = ev[k].scene, j = k;
      while (j < count && ev[j].scene == s) ++j;
      sd[s].evBase = k;
      sd[s].evCount = j - k;
      sd[s].evCursor = k;
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
      HIPCHECK3(hipMalloc((void**)&e->dEvents, sizeof(Event3) * count));
      e->eventCap = count;
    }
    if (count) HIPCHECK3(hipMemcpy(e->dEvents, ev.data(), sizeof(Event3) * count, hipMemcpyHostToDevice));
    e->D.events = e->dEvents;
    HIPCHECK3(hipMemcpy(e->D.sd, sd.data(), sizeof(Scene3D) * e->S, hipMemcpyHostToDevice));
    e->maxEvents = maxGroup;
    drop_graph3(e);
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn3_set_active(sn3_engine* e, const uint8_t* active, uint32_t count) {
  if (!e || !active) return SN_E_ARG;
  if (!e->loaded || count != e->S) return fail3(e, SN_E_ARG, "[aether-3d] activity mask length != scene count");
  try {
    HIPCHECK3(hipStreamSynchronize(e->stream));
    for (uint32_t s = 0; s < count; ++s) e->active[s] = active[s] ? 1 : 0;
    upload_layout3(e);
    drop_graph3(e);
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

int sn3_profile_step(sn3_engine* e, uint32_t count, float out[5]) {
  if (!e || !out) return SN_E_ARG;
  if (!e->loaded) return fail3(e, SN_E_STATE, "[aether-3d] step before load");
  try {
    float acc[5] = {0, 0, 0, 0, 0};
    for (uint32_t k = 0; k < count; ++k) {
      enqueue_substep3(e, e->ev);
      HIPCHECK3(hipEventSynchronize(e->ev[4]));
      for (int m = 0; m < 4; ++m) {
        float t;
        HIPCHECK3(hipEventElapsedTime(&t, e->ev[m], e->ev[m + 1]));
        acc[1 + m] += t;
        acc[0] += t;
      }
    }
    for (int k = 0; k < 5; ++k) out[k] = acc[k];
  } catch (const DevError3& d) {
    return fail3(e, SN_E_DEVICE, d.msg);
  }
  return SN_OK;
}

}  // extern "C"

// Material passes of the 2D substep: friction and the bond gather,
// applyFric, containment, finalize, containStress, bond formation and the
// bond-folded state hash (src/sim.wgsl:1673-2235, 2708-3097,
// src/hash/statehash.wgsl:83-159).  Included by sim2d.hip inside its
// anonymous namespace, after the shared helpers, the hash reduction and the
// water-path kernels it builds on; not a standalone header.
#pragma once

// ========================================================= material passes
// Worlds holding rows with friction or bonds (gen = 1) run the reference's
// remaining passes: friction with the bond gather and applyFric in every
// iteration (src/sim.wgsl:1673-1950), contain + containOverflow before
// finalize (:1973-2235), containStress after thermal (:2742-2771), and on
// formation substeps markCongealNbr + bondForm1 + bondForm2 (:2714-2730,
// :2794-3097); the state hash then folds in the bond rows.  Rigid bodies
// are refused at load, so FLAG_RIGID never occurs here.
constexpr unsigned FPRE = 5, RPRE = 3, BPRE = 5;
constexpr int32_t D2_FLOOR = 262144;
constexpr int32_t FR_TERM = 16777216, BD_TERM = 16777216;
constexpr uint32_t BOND_CAP = 48;
constexpr int32_t BOND_R2 = 17623204, HARD_BREAK2 = 26214400, L0_MAX = 5120, L0_MIN = 512;
constexpr uint32_t BOND_J_MASK = 0x0001ffffu, BOND_L0_SH = 17, BOND_L0_MASK = 0x3ffe0000u,
                   BOND_TOMB = 0x40000000u;
constexpr uint32_t BOND_SALT = 0x5bd1e995u, BOND_SALT2 = 0x2f5b3c17u;
constexpr int32_t SEAL_SEG_MAX2 = 8820900, SEAL_SEG_MIN2 = 670761, SEAL_KEEP4 = 717, SEAL_REACH = 95027,
                  SEAL_STRESS4 = 2560;
constexpr uint32_t MAX_PARTICLES_BONDED = 131071;   // src/engine.js:261 (17-bit partner field)

__device__ __forceinline__ bool seal_liquid(const int32_t* m) {
  return m[M_BONDK] == 0 && m[M_FRIC] == 0 && m[M_GRAV] > 0 && m[M_PHANTOM] == 0;
}
// bond_k (src/sim.wgsl:1655-1663): stiffness at temperature T.
__device__ __forceinline__ int32_t bond_k(const int32_t* m, int32_t T, Dbg& g) {
  if (m[M_BONDK] == 0) return 0;
  if (m[M_RAMPSIGN] == 0) return m[M_BONDK];
  const int32_t span = fp_sub_sat(m[M_RAMPHI], m[M_RAMPLO], g);
  const int32_t dT = wclamp(fp_sub_sat(T, m[M_RAMPLO], g), 0, span);
  int32_t u = clampi(fp_mul_s(dT, m[M_RAMPINV], 16, g), 0, 65536);
  if (m[M_RAMPSIGN] < 0) u = 65536 - u;
  return fp_mul_s(m[M_BONDK], u, 16, g);
}
__device__ __forceinline__ uint32_t* bond_row(const Dev& D, uint32_t* base, const SceneP& P, uint32_t slot) {
  return base + (size_t)(P.partBase + slot) * BOND_CAP;
}
__device__ __forceinline__ int32_t sq2(int32_t x, int32_t y) { return wadd(wmul(x, x), wmul(y, y)); }

// friction (src/sim.wgsl:1673-1934) for fluid slot p: tangential Coulomb
// correction against every neighbour, then the bond springs of p's row
// (breaking ones are tombstoned in place).  FR_L lanes share one particle:
// lane `sub` takes neighbour entries and bond words k = sub, sub + FR_L, ...
// and the lanes' wrapping sums are combined afterwards, which is exact (u32
// sums are order-free, and each lane writes only its own bond words).  A
// row is dense — every word after its first zero is zero — so a lane may
// stop at the first zero it meets.
constexpr uint32_t FR_L = 4;

struct FricSums {
  int32_t fx, fy, bx, by;
};

template <bool BND>
__device__ FricSums friction_part(const Dev& D, const SceneP& P, uint32_t p, uint32_t sub, Dbg& g) {
  FricSums r{0, 0, 0, 0};
  const int4 bi = D.ptY[p];
  const int32_t* Mi = MAT(D, (uint32_t)bi.x);
  const int32_t fi = Mi[M_FRIC];
  if (fi == 0 && Mi[M_BONDK] == 0) return r;   // dp = (0, 0): the zero-work path
  const int2 pi = D.pred[p];
  if (fi != 0) {
    const int4 ai = D.psY[p];
    const int2 dxi = make_int2(fp_sub_sat(pi.x, ai.x, g), fp_sub_sat(pi.y, ai.y, g));
    const uint32_t cnt = D.nbrN[p] & 0x1fffffffu;
    for (uint32_t k = sub; k < cnt; k += FR_L) {
      const uint32_t e = D.nbr[(size_t)k * D.NS + p];
      const uint32_t q = e & ~NBR_BND;
      const int2 pj = D.pred[q];
      const int2 d = pdelta<BND>(pi, pj, g);
      const int32_t id = lut_idx<BND>(d, g);
      if (id < 0) continue;
      const int32_t d4x = fp_shr_rne(d.x, FPRE), d4y = fp_shr_rne(d.y, FPRE);
      const int32_t d2 = sq2(d4x, d4y);
      if (d2 < D2_FLOOR) continue;
      int2 rel = dxi;
      int32_t mu = fi;
      if (!(e & NBR_BND)) {
        const int4 aj = D.psY[q];
        rel = make_int2(fp_sub_sat(dxi.x, fp_sub_sat(pj.x, aj.x, g), g),
                        fp_sub_sat(dxi.y, fp_sub_sat(pj.y, aj.y, g), g));
        mu = mat_mix(fi, MAT(D, (uint32_t)D.ptY[q].x)[M_FRIC]);
      }
      if (mu == 0) continue;
      const int32_t r4x = fp_shr_rne(rel.x, RPRE), r4y = fp_shr_rne(rel.y, RPRE);
      const int32_t dot = wadd(wmul(d4x, r4x), wmul(d4y, r4y));
      const int32_t qq = fp_divshift(dot, d2, 16, g);
      const int32_t tx = fp_sub_sat(r4x, fp_mul_s(d4x, qq, 16, g), g);
      const int32_t ty = fp_sub_sat(r4y, fp_mul_s(d4y, qq, 16, g), g);
      const int32_t wq = fp_shr_rne(D.lut4[id].x, 4);
      const int32_t sx = fp_mul_s(fp_mul_s(tx, wq, 16 - ACC - RPRE, g), mu, 16, g);
      const int32_t sy = fp_mul_s(fp_mul_s(ty, wq, 16 - ACC - RPRE, g), mu, 16, g);
      r.fx = wsub(r.fx, clampi(sx, -FR_TERM, FR_TERM));
      r.fy = wsub(r.fy, clampi(sy, -FR_TERM, FR_TERM));
    }
  }
  if (Mi[M_BONDK] != 0 && ((uint32_t)bi.y & FLAG_BURST) == 0u) {
    const int32_t ki = bond_k(Mi, bi.z, g);
    const int32_t li = D.lam[p];
    uint32_t* row = bond_row(D, D.bond, P, D.idY[p]);
    for (uint32_t k = sub; k < BOND_CAP; k += FR_L) {
      const uint32_t w = row[k];
      if (w == 0u) break;
      if (w & BOND_TOMB) continue;
      const uint32_t q = D.posOf[P.partBase + (w & BOND_J_MASK)];
      const int4 bj = D.ptY[q];
      const int32_t* Mj = MAT(D, (uint32_t)bj.x);
      if (((uint32_t)bj.y & FLAG_BURST) != 0u || Mj[M_BONDK] == 0) { row[k] = w | BOND_TOMB; continue; }
      const int32_t L0 = (int32_t)((w >> BOND_L0_SH) & 0x1fffu);
      const int32_t L02 = wmul(L0, L0);
      const int2 d = pdelta<BND>(pi, D.pred[q], g);
      const int32_t d4x = fp_shr_rne(d.x, BPRE), d4y = fp_shr_rne(d.y, BPRE);
      if (abs(d4x) > L0_MAX || abs(d4y) > L0_MAX) { row[k] = w | BOND_TOMB; continue; }
      const int32_t d2 = sq2(d4x, d4y);
      if (d2 >= HARD_BREAK2) { row[k] = w | BOND_TOMB; continue; }
      int32_t mb = mini(Mi[M_BONDBREAK], Mj[M_BONDBREAK]);
      const int32_t mub = mini(Mi[M_BONDMU], Mj[M_BONDMU]);
      if (mb != 0 && mub != 0) {
        const int32_t lp = fp_neg_sat(fp_add_sat(li, D.lam[q], g), g);
        const int32_t Pq = lp > 0 ? fp_shr_rne(lp, 4) : 0;
        mb = wadd(wadd(65536, wsub(mb, 65536)), fp_mul_s(Pq, mub, 16, g));
      }
      if (mb != 0 && d2 > fp_mul_s(L02, mb, 16, g)) { row[k] = w | BOND_TOMB; continue; }
      if (d2 < D2_FLOOR) continue;
      const int32_t kp = mini(ki, bond_k(Mj, bj.z, g));
      if (kp == 0) continue;
      const int32_t C = fp_sub_sat(d2, L02, g);
      const int32_t qq = fp_divshift(C, wmul(d2, 4), 16, g);
      const int32_t qk = fp_mul_s(qq, kp, 16, g);
      r.bx = wsub(r.bx, clampi(fp_mul_s(d4x, qk, 16 - BPRE - ACC, g), -BD_TERM, BD_TERM));
      r.by = wsub(r.by, clampi(fp_mul_s(d4y, qk, 16 - BPRE - ACC, g), -BD_TERM, BD_TERM));
    }
  }
  return r;
}

// Launched over FR_L x (the per-particle kernels' blocks): logical block lb
// serves 256 / FR_L slots of 256-slot block blockMap[lb / FR_L].
__global__ __launch_bounds__(256) void k_friction(Dev D) {
  const uint32_t lb = xcd_block();
  const uint32_t p = D.blockMap[lb / FR_L] * 256u + (lb % FR_L) * (256u / FR_L) + threadIdx.x / FR_L;
  const uint32_t sub = threadIdx.x % FR_L;
  const uint32_t s = p < D.N ? D.sidOf[p] : INACTIVE;
  Dbg g;
  const bool bnd = wave_bounded(D, s);
  const bool act = s != INACTIVE && (D.info[p] & INF_FLUID);
  FricSums r{0, 0, 0, 0};
  if (act) r = bnd ? friction_part<true>(D, D.sp[s], p, sub, g) : friction_part<false>(D, D.sp[s], p, sub, g);
#pragma unroll
  for (uint32_t off = 1; off < FR_L; off <<= 1) {
    r.fx = wadd(r.fx, __shfl_xor(r.fx, (int)off));
    r.fy = wadd(r.fy, __shfl_xor(r.fy, (int)off));
    r.bx = wadd(r.bx, __shfl_xor(r.bx, (int)off));
    r.by = wadd(r.by, __shfl_xor(r.by, (int)off));
  }
  if (act && sub == 0u) D.dp[p] = make_int2(fp_add_sat(r.fx, r.bx, g), fp_add_sat(r.fy, r.by, g));
  flush_dbg(D, s, g, 0);
}

// applyFric (src/sim.wgsl:1945-1950) = applyDp_at on every fluid slot.
__global__ __launch_bounds__(256) void k_apply_fric(Dev D) {
  const uint32_t p = slot_block(D) * 256u + threadIdx.x;
  const uint32_t s = p < D.N ? D.sidOf[p] : INACTIVE;
  Dbg g;
  if (s != INACTIVE && (D.info[p] & INF_FLUID)) {
    const int4 b = D.ptY[p];
    if (((uint32_t)b.y & (FLAG_RIGID | FLAG_ANCHOR)) == 0u) {
      const int32_t* M = MAT(D, (uint32_t)b.x);
      const int32_t dm = M[M_DPMAX], dmN = fp_neg_sat(dm, g);
      const int2 dp = D.dp[p];
      const int2 pr = D.pred[p];
      const int32_t ddx = clampi(fp_shr_rne(dp.x, ACC), dmN, dm);
      const int32_t ddy = clampi(fp_shr_rne(dp.y, ACC), dmN, dm);
      D.pred[p] = wall_clamp(D.sp[s], make_int2(fp_add_sat(pr.x, ddx, g), fp_add_sat(pr.y, ddy, g)));
    }
  }
  flush_dbg(D, s, g, 0);
}

// ---- containment helpers (src/sim.wgsl:1994-1999, :2085-2135)
__device__ __forceinline__ int32_t seal_cross_sign(int32_t ax, int32_t ay, int32_t bx, int32_t by) {
  // sign of the wrapped 64-bit a.x*b.y - a.y*b.x
  const uint64_t d = (uint64_t)((int64_t)ax * by) - (uint64_t)((int64_t)ay * bx);
  if (d == 0) return 0;
  return (d >> 63) ? -1 : 1;
}
// Push pr out to the keep-out radius of seal particle pj.
__device__ __forceinline__ void seal_keep_push(int2 pj, int2& pr, Dbg& g) {
  const int32_t dx4 = fp_shr_rne(fp_sub_sat(pr.x, pj.x, g), BPRE);
  const int32_t dy4 = fp_shr_rne(fp_sub_sat(pr.y, pj.y, g), BPRE);
  const int32_t d2 = sq2(dx4, dy4);
  if (d2 > 0 && d2 < SEAL_KEEP4 * SEAL_KEEP4) {
    const int32_t d4 = (int32_t)isqrt32((uint32_t)d2);
    if (d4 > 0) {
      const int32_t ux = fp_div_s(dx4, d4, 16, g), uy = fp_div_s(dy4, d4, 16, g);
      const int32_t amt = (SEAL_KEEP4 - d4) << BPRE;
      pr.x = fp_add_sat(pr.x, fp_mul_s(ux, amt, 16, g), g);
      pr.y = fp_add_sat(pr.y, fp_mul_s(uy, amt, 16, g), g);
    }
  }
}
// Revert the move pos -> pr if it crosses the seal segment pj-pk.
__device__ __forceinline__ void seal_segment_test(int2 pj, int2 pk, int2 pos, int2& pr, Dbg& g) {
  const int32_t ex4 = fp_shr_rne(fp_sub_sat(pk.x, pj.x, g), BPRE);
  const int32_t ey4 = fp_shr_rne(fp_sub_sat(pk.y, pj.y, g), BPRE);
  const int32_t e2 = sq2(ex4, ey4);
  if (e2 > SEAL_SEG_MAX2 || e2 < SEAL_SEG_MIN2) return;
  if (mini(pj.x, pk.x) > maxi(pos.x, pr.x) || maxi(pj.x, pk.x) < mini(pos.x, pr.x) ||
      mini(pj.y, pk.y) > maxi(pos.y, pr.y) || maxi(pj.y, pk.y) < mini(pos.y, pr.y))
    return;
  const int32_t ex = fp_sub_sat(pk.x, pj.x, g), ey = fp_sub_sat(pk.y, pj.y, g);
  const int32_t s0 = seal_cross_sign(ex, ey, fp_sub_sat(pos.x, pj.x, g), fp_sub_sat(pos.y, pj.y, g));
  const int32_t s1 = seal_cross_sign(ex, ey, fp_sub_sat(pr.x, pj.x, g), fp_sub_sat(pr.y, pj.y, g));
  if (s0 == 0 || s0 == s1) return;
  const int32_t mx = fp_sub_sat(pr.x, pos.x, g), my = fp_sub_sat(pr.y, pos.y, g);
  const int32_t t0 = seal_cross_sign(mx, my, fp_sub_sat(pj.x, pos.x, g), fp_sub_sat(pj.y, pos.y, g));
  const int32_t t1 = seal_cross_sign(mx, my, fp_sub_sat(pk.x, pos.x, g), fp_sub_sat(pk.y, pos.y, g));
  if (t0 == t1 && t0 != 0) return;
  pr = pos;
}

constexpr int64_t SEAL_NEAR = 98304;   // 1.5 wu: longer than any kept seal segment (1.4504 wu)

// containOverflow's candidate sequence: up to three contiguous ranges of
// sorted positions (one per grid row of the 3x3 block), filtered to the
// seal-class fluid slots whose pred lies in the reach box.
struct SealBox {
  int32_t loX, hiX, loY, hiY;
  uint32_t lo[3], hi[3];
  uint32_t nr;
};
struct SealCursor {
  uint32_t r, t;   // range index, position
};
__device__ __forceinline__ SealCursor seal_begin(const SealBox& B) {
  SealCursor c;
  c.r = 0;
  c.t = B.nr ? B.lo[0] : 0u;
  return c;
}
// Advance c to the first qualifying position at or after it; false at the end.
__device__ __forceinline__ bool seal_next(const Dev& D, const SceneP& P, const SealBox& B, SealCursor& c) {
  while (c.r < B.nr) {
    if (c.t >= B.hi[c.r]) {
      if (++c.r < B.nr) c.t = B.lo[c.r];
      continue;
    }
    const uint32_t t = c.t;
    if (D.idY[t] < P.nFluid && seal_class(MAT(D, (uint32_t)D.ptY[t].x))) {
      const int2 pj = D.pred[t];
      if (!(pj.x < B.loX || pj.x > B.hiX || pj.y < B.loY || pj.y > B.hiY)) return true;
    }
    c.t++;
  }
  return false;
}

// contain (src/sim.wgsl:2056-2136) for a list below MAXNBR and
// containOverflow (:2147-2235) for a full one, in one launch: both write only
// the pred of a non-seal slot and read only the preds of seal-class slots,
// which neither moves.
__device__ void contain_at(const Dev& D, const SceneP& P, uint32_t p, Dbg& g) {
  const uint32_t nn = D.nbrN[p];
  if ((nn & NBR_SEAL_BIT) == 0u) return;
  const int4 bi = D.ptY[p];
  const int32_t* Mi = MAT(D, (uint32_t)bi.x);
  if (seal_class(Mi) || Mi[M_PHANTOM] != 0) return;
  if (((uint32_t)bi.y & (FLAG_ANCHOR | FLAG_RIGID)) != 0u) return;
  const int4 ai = D.psY[p];
  const int2 pos = make_int2(ai.x, ai.y);
  int2 pr = D.pred[p];
  const bool liquidI = seal_liquid(Mi);
  const uint32_t cnt = nn & 0x1fffffffu;
  if (cnt < MAXNBR && P.bounded) {
    // The crossing tests only ever set pr back to pos, and once pr == pos no
    // test fires again (s0 == s1), so the pass reverts exactly when some
    // pair of seal neighbours, tested against the pushed pr, fires: an OR
    // over the pairs.  A pair fires only if its segment's box meets the
    // move's box, and a kept segment is at most 1.4504 wu long (e2 <=
    // SEAL_SEG_MAX2), so an end farther than SEAL_NEAR from the move's box
    // on either axis rules out every pair it is in.  Positions of a bounded
    // scene cannot saturate, so skipped tests would have counted no event.
    uint64_t sm = 0;   // neighbour entries that are seal-class fluid slots
    for (uint32_t k = 0; k < cnt; ++k) {
      const uint32_t e = D.nbr[(size_t)k * D.NS + p];
      if (e & NBR_BND) continue;
      if (!((D.sealRows >> ((uint32_t)D.ptY[e].x & 63u)) & 1ull)) continue;
      sm |= 1ull << k;
      seal_keep_push(D.pred[e], pr, g);
    }
    if (sm != 0u) {
      const int64_t bx0 = (int64_t)mini(pos.x, pr.x) - SEAL_NEAR, bx1 = (int64_t)maxi(pos.x, pr.x) + SEAL_NEAR;
      const int64_t by0 = (int64_t)mini(pos.y, pr.y) - SEAL_NEAR, by1 = (int64_t)maxi(pos.y, pr.y) + SEAL_NEAR;
      uint64_t live = 0;
      for (uint64_t m = sm; m; m &= m - 1u) {
        const uint32_t k = (uint32_t)__builtin_ctzll(m);
        const int2 pj = D.pred[D.nbr[(size_t)k * D.NS + p]];
        if (pj.x >= bx0 && pj.x <= bx1 && pj.y >= by0 && pj.y <= by1) live |= 1ull << k;
      }
      // once reverted (pr == pos) nothing fires again: stop there
      for (uint64_t m = live; m != 0u && (pr.x != pos.x || pr.y != pos.y); m &= m - 1u) {
        const uint32_t k = (uint32_t)__builtin_ctzll(m);
        const uint32_t e = D.nbr[(size_t)k * D.NS + p];
        const int2 pj = D.pred[e];
        const bool modeJ = liquidI && MAT(D, (uint32_t)D.ptY[e].x)[M_SEALMODE] != 0;
        for (uint64_t m2 = live & ~((2ull << k) - 1ull); m2; m2 &= m2 - 1u) {
          const uint32_t e2 = D.nbr[(size_t)__builtin_ctzll(m2) * D.NS + p];
          if (modeJ && MAT(D, (uint32_t)D.ptY[e2].x)[M_SEALMODE] != 0) continue;
          seal_segment_test(pj, D.pred[e2], pos, pr, g);
        }
      }
    }
  } else if (cnt < MAXNBR) {
    bool touched = false;
    for (uint32_t k = 0; k < cnt; ++k) {
      const uint32_t e = D.nbr[(size_t)k * D.NS + p];
      if (e & NBR_BND) continue;
      if (!seal_class(MAT(D, (uint32_t)D.ptY[e].x))) continue;
      touched = true;
      seal_keep_push(D.pred[e], pr, g);
    }
    if (touched) {
      for (uint32_t k = 0; k < cnt; ++k) {
        const uint32_t e = D.nbr[(size_t)k * D.NS + p];
        if (e & NBR_BND) continue;
        const int32_t* Mj = MAT(D, (uint32_t)D.ptY[e].x);
        if (!seal_class(Mj)) continue;
        const int2 pj = D.pred[e];
        for (uint32_t k2 = k + 1; k2 < cnt; ++k2) {
          const uint32_t e2 = D.nbr[(size_t)k2 * D.NS + p];
          if (e2 & NBR_BND) continue;
          const int32_t* Mk = MAT(D, (uint32_t)D.ptY[e2].x);
          if (!seal_class(Mk)) continue;
          if (liquidI && Mj[M_SEALMODE] != 0 && Mk[M_SEALMODE] != 0) continue;
          seal_segment_test(pj, D.pred[e2], pos, pr, g);
        }
      }
    }
  } else {
    // The seal-class slots of the 3x3 cells around pred, in canonical order,
    // within SEAL_REACH of the swept box; the first 48 are used.  The cells
    // gx-1..gx+1 of one grid row are consecutive in the cell-sorted order, so
    // the candidates are up to three contiguous position ranges, enumerated
    // again for each use instead of stored (no private array, no scratch).
    SealBox B;
    B.loX = fp_sub_sat(mini(pos.x, pr.x), SEAL_REACH, g);
    B.hiX = fp_add_sat(maxi(pos.x, pr.x), SEAL_REACH, g);
    B.loY = fp_sub_sat(mini(pos.y, pr.y), SEAL_REACH, g);
    B.hiY = fp_add_sat(maxi(pos.y, pr.y), SEAL_REACH, g);
    B.nr = 0;
    const CC cc = cell_coord(P, pr);
    const int32_t gx0 = maxi(cc.x - 1, 0), gx1 = mini(cc.x + 1, (int32_t)P.gridW - 1);
    for (int32_t gy = cc.y - 1; gy <= cc.y + 1; ++gy) {
      if (gy < 0 || gy >= (int32_t)P.gridH) continue;
      const uint32_t c0 = P.cellBase + (uint32_t)gy * P.gridW + (uint32_t)gx0;
      const uint32_t c1 = P.cellBase + (uint32_t)gy * P.gridW + (uint32_t)gx1;
      B.lo[B.nr] = (uint32_t)((int32_t)D.cellStart[c0] + P.posShift);
      B.hi[B.nr] = (uint32_t)((int32_t)D.cellStart[c1 + 1] + P.posShift);
      B.nr++;
    }
    SealCursor a = seal_begin(B);
    for (uint32_t k = 0; k < 48u && seal_next(D, P, B, a); ++k, ++a.t) seal_keep_push(D.pred[a.t], pr, g);
    a = seal_begin(B);
    for (uint32_t k = 0; k < 48u && seal_next(D, P, B, a); ++k) {
      const int32_t* Mj = MAT(D, (uint32_t)D.ptY[a.t].x);
      const int2 pj = D.pred[a.t];
      SealCursor b = a;
      b.t++;
      for (uint32_t k2 = k + 1; k2 < 48u && seal_next(D, P, B, b); ++k2, ++b.t) {
        const int32_t* Mk = MAT(D, (uint32_t)D.ptY[b.t].x);
        if (liquidI && Mj[M_SEALMODE] != 0 && Mk[M_SEALMODE] != 0) continue;
        seal_segment_test(pj, D.pred[b.t], pos, pr, g);
      }
      a.t++;
    }
  }
  D.pred[p] = pr;
}

// contain + containOverflow, then finalize (src/sim.wgsl:2245-2265) in a
// second launch: finalize reads only the slot's own pred, but contain reads
// seal preds that a fused finalize would not change either way; the split
// keeps the finalize of k_apply's last colour pass as the one definition.
__global__ __launch_bounds__(256) void k_contain(Dev D) {
  const uint32_t p = slot_block(D) * 256u + threadIdx.x;
  const uint32_t s = p < D.N ? D.sidOf[p] : INACTIVE;
  Dbg g;
  if (s != INACTIVE && (D.info[p] & INF_FLUID)) contain_at(D, D.sp[s], p, g);
  flush_dbg(D, s, g, 0);
}

__global__ __launch_bounds__(256) void k_finalize(Dev D) {
  const uint32_t p = slot_block(D) * 256u + threadIdx.x;
  const uint32_t s = p < D.N ? D.sidOf[p] : INACTIVE;
  Dbg g;
  if (s != INACTIVE && (D.info[p] & INF_FLUID)) {
    const SceneP& P = D.sp[s];
    const int2 pr = D.pred[p];
    const int4 a = D.psY[p];
    const int32_t vmaxN = fp_neg_sat(P.vmax, g);
    const int32_t vx = clampi(fp_sub_sat(pr.x, a.x, g), vmaxN, P.vmax);
    const int32_t vy = clampi(fp_sub_sat(pr.y, a.y, g), vmaxN, P.vmax);
    const int32_t rho = D.rho[p];
    const int32_t ri = clampi(fp_divshift(RHO0, maxi(rho, RHO_FLOOR_RINV), 16, g), 0, RINV_MAX);
    D.vtmp[p] = make_int2(vx, vy);
    D.rinv[p] = ri;
    D.vr4[p] = make_int4(vx, vy, rho, ri);   // xnc's gather record
  }
  flush_dbg(D, s, g, 0);
}

// containStress (src/sim.wgsl:2742-2771): a held seal particle crushed by
// non-seal neighbours bursts.  Writes only its own flags word; reads the
// positions and material ids of others, which no slot changes here.
__global__ __launch_bounds__(256) void k_contain_stress(Dev D) {
  const uint32_t p = slot_block(D) * 256u + threadIdx.x;
  const uint32_t s = p < D.N ? D.sidOf[p] : INACTIVE;
  Dbg g;
  if (s != INACTIVE && (D.info[p] & INF_FLUID)) {
    const int4 b = D.pt[p];
    const int32_t* Mi = MAT(D, (uint32_t)b.x);
    const uint32_t fl = (uint32_t)b.y;
    if (seal_class(Mi) && Mi[M_BURSTP] > 0 && (fl & (FLAG_ANCHOR | FLAG_RIGID)) != 0u) {
      const int4 a = D.ps[p];
      const uint32_t cnt = D.nbrN[p] & 0x1fffffffu;
      int32_t stress = 0;
      for (uint32_t k = 0; k < cnt; ++k) {
        const uint32_t e = D.nbr[(size_t)k * D.NS + p];
        if (e & NBR_BND) continue;
        const int32_t* Mj = MAT(D, (uint32_t)D.pt[e].x);
        if (seal_class(Mj) || Mj[M_PHANTOM] != 0) continue;
        const int4 aj = D.ps[e];
        const int32_t dx4 = fp_shr_rne(fp_sub_sat(a.x, aj.x, g), BPRE);
        const int32_t dy4 = fp_shr_rne(fp_sub_sat(a.y, aj.y, g), BPRE);
        const int32_t d2 = sq2(dx4, dy4);
        if (d2 >= SEAL_STRESS4 * SEAL_STRESS4) continue;
        stress = wadd(stress, SEAL_STRESS4 - (int32_t)isqrt32((uint32_t)d2));
      }
      if (wmul(stress, 1 << BPRE) > Mi[M_BURSTP])
        reinterpret_cast<int32_t*>(D.pt + p)[1] =
            (int32_t)((fl & ~(FLAG_ANCHOR | FLAG_RIGID | WELD_TAG_MASK)) | FLAG_BURST);
    }
  }
  flush_dbg(D, s, g, 0);
}

__device__ __forceinline__ bool formation_substep(const SceneD& SD) {
  return ((SD.substep % BOND_PERIOD) == 0u) || SD.bondDirty;
}

// markCongealNbr (src/sim.wgsl:2714-2730) folded into bondForm1
// (:2794-3027): the congeal-near bit is only read by bondForm1 of the same
// slot, so it is computed where it is used.  Candidate rows: the live
// bonds, then new partners in canonical neighbour order (weld tags,
// congealing, or reform under the velocity and pressure gates).
__global__ __launch_bounds__(256) void k_bond_form1(Dev D) {
  const uint32_t p = slot_block(D) * 256u + threadIdx.x;
  const uint32_t s = p < D.N ? D.sidOf[p] : INACTIVE;
  if (s == INACTIVE || !(D.info[p] & INF_FLUID) || !formation_substep(D.sd[s])) return;
  Dbg g;
  uint32_t overflow = 0;
  const SceneP& P = D.sp[s];
  const uint32_t myid = D.id[p];
  uint32_t* cand = bond_row(D, D.bondCand, P, myid);
  const int4 a = D.ps[p], b = D.pt[p];
  const uint32_t fl = (uint32_t)b.y, mi = (uint32_t)b.x;
  const bool anchoredI = (fl & FLAG_ANCHOR) != 0u;
  const int32_t* Mi = MAT(D, mi);
  uint32_t cn = 0;
  if (Mi[M_BONDK] != 0 && Mi[M_BONDGROUP] >= 0) {
    const uint32_t* row = bond_row(D, D.bond, P, myid);
    for (uint32_t k = 0; k < BOND_CAP; ++k) {
      const uint32_t w = row[k];
      if (w == 0u) break;
      if (w & BOND_TOMB) continue;
      cand[cn++] = w;   // a row holds at most BOND_CAP words, so cn < BOND_CAP here
    }
    const uint32_t tagI = fl & WELD_TAG_MASK;
    const int32_t ki = bond_k(Mi, b.z, g);
    const bool canReform = !anchoredI && Mi[M_BONDREFORM] != 0 && ki != 0;
    const bool fuseI = fuse_row(Mi, mi);
    const bool congealI = (fl & FLAG_CONGEAL) != 0u && fuseI;
    const uint32_t cnt = D.nbrN[p] & 0x1fffffffu;
    bool congealNear = false;
    if (fuseI) {
      for (uint32_t k = 0; k < cnt; ++k) {
        const uint32_t e = D.nbr[(size_t)k * D.NS + p];
        if (!(e & NBR_BND) && ((uint32_t)D.pt[e].y & FLAG_CONGEAL) != 0u) { congealNear = true; break; }
      }
    }
    if (canReform || (!anchoredI && tagI != 0u) || congealI || congealNear) {
      for (uint32_t k = 0; k < cnt; ++k) {
        if (cn >= BOND_CAP) { overflow++; break; }
        const uint32_t e = D.nbr[(size_t)k * D.NS + p];
        if (e & NBR_BND) continue;                      // no bonds to boundary
        const int4 bj = D.pt[e];
        const uint32_t mj = (uint32_t)bj.x, fJ = (uint32_t)bj.y;
        const int32_t* Mj = MAT(D, mj);
        const uint32_t tagJ = fJ & WELD_TAG_MASK;
        const bool weld = !anchoredI && tagI != 0u && tagI == tagJ;
        const bool congealJ = (fJ & FLAG_CONGEAL) != 0u && fuse_row(Mj, mj);
        const bool congeal = (congealI || congealJ) && !weld && mj == mi && (fJ & FLAG_RIGID) == 0u;
        if (!congeal && Mj[M_BONDGROUP] != Mi[M_BONDGROUP]) continue;
        if (!weld && !congeal) {
          if (!canReform) continue;
          if (bond_k(Mj, bj.z, g) == 0) continue;
        }
        const uint32_t j = D.id[e];
        bool dup = false;
        for (uint32_t q = 0; q < cn; ++q)
          if ((cand[q] & BOND_J_MASK) == j) { dup = true; break; }
        if (dup) continue;
        const int4 aj = D.ps[e];
        const int32_t d4x = fp_shr_rne(fp_sub_sat(a.x, aj.x, g), BPRE);
        const int32_t d4y = fp_shr_rne(fp_sub_sat(a.y, aj.y, g), BPRE);
        if (abs(d4x) > L0_MAX || abs(d4y) > L0_MAX) continue;
        const int32_t d2 = sq2(d4x, d4y);
        if (d2 < D2_FLOOR || d2 > BOND_R2) continue;
        if (!weld && !congeal) {
          const int32_t vl4 = fp_shr_rne(mini(Mi[M_BONDFORMV], Mj[M_BONDFORMV]), RPRE);
          const int32_t r4x = fp_shr_rne(fp_sub_sat(a.z, aj.z, g), RPRE);
          const int32_t r4y = fp_shr_rne(fp_sub_sat(a.w, aj.w, g), RPRE);
          if (abs(r4x) > vl4 || abs(r4y) > vl4) continue;
          if (sq2(r4x, r4y) > wmul(vl4, vl4)) continue;
          const int32_t pl = mini(Mi[M_BONDFORMP], Mj[M_BONDFORMP]);
          if (pl != 0 && (fp_neg_sat(D.lam[p], g) < pl || fp_neg_sat(D.lam[e], g) < pl)) continue;
        }
        const int32_t L0 = (int32_t)isqrt32((uint32_t)d2);
        if (L0 < L0_MIN || L0 > L0_MAX) continue;
        cand[cn++] = (((uint32_t)L0 & 0x1fffu) << BOND_L0_SH) | (j & BOND_J_MASK);
      }
    }
  }
  for (uint32_t k = cn; k < BOND_CAP; ++k) cand[k] = 0u;
  if (D.dbgOn) {
    Dbg z = g;
    if (z.mul | z.div | z.div0 | z.add) {
      if (z.mul) atomicAdd(&D.sd[s].dbg[0], z.mul);
      if (z.div) atomicAdd(&D.sd[s].dbg[1], z.div);
      if (z.div0) atomicAdd(&D.sd[s].dbg[2], z.div0);
      if (z.add) atomicAdd(&D.sd[s].dbg[5], z.add);
    }
    if (overflow) atomicAdd(&D.sd[s].dbg[8], overflow);
  }
}

// bondForm2 (src/sim.wgsl:3039-3097): keep the candidates whose partner
// proposed the same rest length back, then the formation flag update.
__global__ __launch_bounds__(256) void k_bond_form2(Dev D) {
  const uint32_t p = slot_block(D) * 256u + threadIdx.x;
  const uint32_t s = p < D.N ? D.sidOf[p] : INACTIVE;
  if (s == INACTIVE || !(D.info[p] & INF_FLUID) || !formation_substep(D.sd[s])) return;
  const SceneP& P = D.sp[s];
  const uint32_t myid = D.id[p];
  const uint32_t* cand = bond_row(D, D.bondCand, P, myid);
  uint32_t* row = bond_row(D, D.bond, P, myid);
  uint32_t n = 0;
  for (uint32_t k = 0; k < BOND_CAP; ++k) {
    const uint32_t w = cand[k];
    if (w == 0u) break;
    const uint32_t* other = bond_row(D, D.bondCand, P, w & BOND_J_MASK);
    bool ok = false;
    for (uint32_t q = 0; q < BOND_CAP; ++q) {
      const uint32_t v = other[q];
      if (v == 0u) break;
      if ((v & BOND_J_MASK) == myid) { ok = (v & BOND_L0_MASK) == (w & BOND_L0_MASK); break; }
    }
    if (ok) row[n++] = w & ~BOND_TOMB;
  }
  for (uint32_t k = n; k < BOND_CAP; ++k) row[k] = 0u;
  uint32_t f = (uint32_t)D.pt[p].y & ~FLAG_CONGEAL;
  if ((f & FLAG_RIGID) == 0u && (f & WELD_TAG_MASK) != P.weldHold) f &= ~WELD_TAG_MASK;
  reinterpret_cast<int32_t*>(D.pt + p)[1] = (int32_t)f;
}

// The state hash of a world running the material passes: as in
// k_st_thermal_hash, with each fluid slot's bond row folded in.
__global__ __launch_bounds__(256) void k_hash_gen(Dev D) {
  const uint32_t blk = slot_block(D);
  const uint32_t p = blk * 256u + threadIdx.x;
  const uint32_t s = p < D.N ? D.sidOf[p] : INACTIVE;
  uint32_t hv[HPART_WORDS];
  if (s != INACTIVE) {
    const SceneP& P = D.sp[s];
    const int4 a = D.ps[p], b = D.pt[p];
    const uint32_t myid = D.id[p];
    const uint32_t words[8] = {(uint32_t)a.x, (uint32_t)a.y, (uint32_t)a.z, (uint32_t)a.w,
                               (uint32_t)b.x, (uint32_t)b.y, (uint32_t)b.z, (uint32_t)b.w};
    uint32_t bn = 0, com = 0, scom = 0;
    if (myid < P.nFluid) {
      const uint32_t* row = bond_row(D, D.bond, P, myid);
      for (uint32_t k = 0; k < BOND_CAP; ++k) {
        const uint32_t w = row[k];
        if (w == 0u) break;
        com += fmix32(((w >> 17) & 0x1fffu) ^ BOND_SALT);
        scom += fmix32(w ^ BOND_SALT2);
        bn++;
      }
    }
    hash_terms(words, myid, bn, com, scom, hv);
    hv[8] = D.drop[p];
  } else {
#pragma unroll
    for (int l = 0; l < HPART_WORDS; ++l) hv[l] = 0u;
  }
  hash_reduce(D, blk, s, hv);
}

// Digest of the current state without advancing (digest(), src/engine.js:4453),
// with the bond fold in a world that runs the material passes.
__global__ __launch_bounds__(256) void k_hash_state(Dev D, uint32_t s) {
  const SceneP& P = D.sp[s];
  uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= P.n) return;
  uint32_t p = P.partBase + t;
  int4 a = D.ps[p], b = D.pt[p];
  uint32_t myid = D.id[p];
  uint32_t words[8] = {(uint32_t)a.x, (uint32_t)a.y, (uint32_t)a.z, (uint32_t)a.w,
                       (uint32_t)b.x, (uint32_t)b.y, (uint32_t)b.z, (uint32_t)b.w};
  uint32_t bn = 0, com = 0, scom = 0;
  if (D.gen && myid < P.nFluid) {
    const uint32_t* row = bond_row(D, D.bond, P, myid);
    for (uint32_t k = 0; k < BOND_CAP; ++k) {
      const uint32_t w = row[k];
      if (w == 0u) break;
      com += fmix32(((w >> 17) & 0x1fffu) ^ BOND_SALT);
      scom += fmix32(w ^ BOND_SALT2);
      bn++;
    }
  }
  uint32_t hv[8];
  hash_terms(words, myid, bn, com, scom, hv);
  for (int l = 0; l < 8; ++l) atomicAdd(&D.hacc[s * 8 + l], hv[l]);
}

// Per scene: fold the chain, append this substep's digest to the H ring,
// clear the accumulators, advance the substep counter.
// One wave per scene: fold the scene's block partials (blocks lying wholly
// inside its slot range) into the atomically summed remainder, then the
// chain step, the H ring append and the substep counter.
__global__ __launch_bounds__(64) void k_hash_finish(Dev D, const uint8_t* active) {
  const uint32_t s = blockIdx.x;
  if (!active[s]) return;
  SceneD& SD = D.sd[s];
  const SceneP& P = D.sp[s];
  const uint32_t lane = threadIdx.x;
  const uint32_t b0 = (P.partBase + 255u) / 256u, b1 = (P.partBase + P.n) / 256u;
  uint32_t sum[HPART_WORDS];
#pragma unroll
  for (int l = 0; l < HPART_WORDS; ++l) sum[l] = 0u;
  for (uint32_t b = b0 + lane; b < b1; b += 64u) {
#pragma unroll
    for (int l = 0; l < HPART_WORDS; ++l) sum[l] += D.hpart[(size_t)b * HPART_WORDS + l];
  }
#pragma unroll
  for (int l = 0; l < HPART_WORDS; ++l) sum[l] = wave_sum(sum[l]);
  if (lane != 0u) return;
  uint32_t a[8];
  for (int l = 0; l < 8; ++l) a[l] = SD.acc[l] + sum[l];
  SD.dbg[6] += sum[8];
  SD.flagged = 0u;
  for (int j = 0; j < 8; ++j) SD.chain[j] = mm_step(SD.chain[j], a[j] ^ rotl32(a[(j + 1) & 7], 7));
  uint32_t slot = SD.streamPos % D.ringCap;
  uint32_t* r = D.hring + ((size_t)s * D.ringCap + slot) * 8;
  for (int l = 0; l < 8; ++l) { r[l] = a[l]; SD.acc[l] = 0u; }
  SD.streamPos++;
  bool formation = ((SD.substep % BOND_PERIOD) == 0u) || SD.bondDirty;
  if (formation) SD.bondDirty = 0u;
  SD.substep++;
}

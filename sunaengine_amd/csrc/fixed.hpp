// Q16.16 integer core for gfx950 (and the host, for CPU tests of this header).
//
// Bit-exact restatement of the reference's fixed-point API (src/fixed.wgsl):
// fp_mul_s / i64_shr_rne_sat (:146-170), fp_add_sat / fp_sub_sat / fp_neg_sat /
// fp_abs (:175-200), fp_shr_rne (:205-213), fp_divshift (:285-309), isqrt64
// (:318-332).  The reference builds 64-bit products from 16-bit halves; on
// CDNA4 the exact product is one v_mul_lo_u32 + v_mul_hi_i32 pair, and the
// round-half-to-even shift is folded into one biased 64-bit shift.
//
// Debug events (src/fixed.wgsl:62-77) are counted into a per-thread `Dbg`
// held in registers and flushed once per kernel; the counters are u32 sums,
// so the totals do not depend on execution order.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define SN_HD __host__ __device__ __forceinline__
#else
#define SN_HD inline
#endif

namespace sn {

constexpr int32_t I32_MAX = 2147483647;
constexpr int32_t I32_MIN = -2147483647 - 1;

struct Dbg {
  uint32_t mul = 0;   // dbg[0] multiply/shift saturation
  uint32_t div = 0;   // dbg[1] divide saturation / out-of-contract shift
  uint32_t div0 = 0;  // dbg[2] divide by zero
  uint32_t add = 0;   // dbg[5] add/sub/neg saturation
};

SN_HD int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
SN_HD int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
SN_HD int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
SN_HD int32_t clampi(int32_t x, int32_t lo, int32_t hi) { return x < lo ? lo : (x > hi ? hi : x); }
SN_HD int32_t maxi(int32_t a, int32_t b) { return a > b ? a : b; }
SN_HD int32_t mini(int32_t a, int32_t b) { return a < b ? a : b; }
// WGSL clamp(x, lo, hi) = min(max(x, lo), hi), also when lo > hi (clampi
// assumes lo <= hi, which holds wherever its bounds are constants).
SN_HD int32_t wclamp(int32_t x, int32_t lo, int32_t hi) { return mini(maxi(x, lo), hi); }

// Every operation below is branch-free: the saturated result is selected
// and the debug counter advanced by the overflow predicate, so a wave never
// splits on the (almost always false) saturation test.
SN_HD int32_t sat_of(int32_t sign_src) { return (sign_src >> 31) ^ I32_MAX; }  // I32_MIN if < 0 else I32_MAX

// RNE(v / 2^s) saturated to i32, s in 1..31.  floor-shift of v biased by
// (2^(s-1) - 1 + lsb(floor(v / 2^s))) is exactly round-half-to-even.
SN_HD int32_t shr_rne_sat64(int64_t v, unsigned s, Dbg& g) {
  int64_t odd = (v >> s) & 1;
  int64_t q = (v + ((int64_t)1 << (s - 1)) - 1 + odd) >> s;
  const bool ovf = (int64_t)(int32_t)q != q;
  g.mul += ovf ? 1u : 0u;
  return ovf ? sat_of((int32_t)(v >> 32)) : (int32_t)q;
}

SN_HD int32_t fp_mul_s(int32_t a, int32_t b, unsigned s, Dbg& g) {
  return shr_rne_sat64((int64_t)a * (int64_t)b, s, g);
}

SN_HD int32_t fp_add_sat(int32_t a, int32_t b, Dbg& g) {
  const int32_t s = wadd(a, b);
  const bool ovf = ((a ^ s) & (b ^ s)) < 0;
  g.add += ovf ? 1u : 0u;
  return ovf ? sat_of(a) : s;
}
SN_HD int32_t fp_sub_sat(int32_t a, int32_t b, Dbg& g) {
  const int32_t s = wsub(a, b);
  const bool ovf = ((a ^ b) & (a ^ s)) < 0;
  g.add += ovf ? 1u : 0u;
  return ovf ? sat_of(a) : s;
}
SN_HD int32_t fp_neg_sat(int32_t a, Dbg& g) {
  const bool ovf = a == I32_MIN;
  g.add += ovf ? 1u : 0u;
  return ovf ? I32_MAX : wsub(0, a);
}
SN_HD int32_t fp_abs(int32_t a, Dbg& g) {
  const bool ovf = a == I32_MIN;
  g.add += ovf ? 1u : 0u;
  return ovf ? I32_MAX : (a < 0 ? wsub(0, a) : a);
}

// RNE arithmetic shift of an i32 (|result| <= 2^30, cannot overflow).
SN_HD int32_t fp_shr_rne(int32_t x, unsigned s) {
  int32_t odd = (x >> s) & 1;
  return (int32_t)(((int64_t)x + ((int64_t)1 << (s - 1)) - 1 + odd) >> s);
}

SN_HD int clz32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return x == 0 ? 32 : __clz((int)x);
#else
  return x == 0 ? 32 : __builtin_clz(x);
#endif
}

// floor(2^30 / dn) for dn in [2^15, 2^16): float estimate, then an exact
// integer remainder correction (the estimate is within 1 of the quotient).
SN_HD int32_t recip30(int32_t dn) {
#if defined(__HIP_DEVICE_COMPILE__)
  int32_t q = (int32_t)(1073741824.0f * __builtin_amdgcn_rcpf((float)dn));
#else
  int32_t q = (int32_t)(1073741824.0f / (float)dn);
#endif
  int32_t r = 1073741824 - q * dn;
  if (r < 0) { q -= 1; r += dn; }
  if (r < 0) { q -= 1; r += dn; }
  if (r >= dn) { q += 1; r -= dn; }
  if (r >= dn) { q += 1; }
  return q;
}

// fp_divshift (src/fixed.wgsl:285-309).
SN_HD int32_t fp_divshift(int32_t num, int32_t den, unsigned k, Dbg& g) {
  if (den == 0) {
    g.div0++;
    if (num == 0) return 0;
    return num > 0 ? I32_MAX : I32_MIN;
  }
  int32_t d = den, nn = num;
  if (d < 0) { g.div++; d = fp_abs(d, g); nn = fp_neg_sat(nn, g); }
  int e = 16 - clz32((uint32_t)d);
  int32_t dn = e >= 0 ? (d >> e) : (d << (-e));
  int32_t Rn = recip30(dn);
  int s = 30 + e - (int)k;
  if (s > 31) { g.div++; return 0; }
  if (s < 1) {
    g.div++;
    if (nn == 0) return 0;
    return nn > 0 ? I32_MAX : I32_MIN;
  }
  return shr_rne_sat64((int64_t)nn * (int64_t)Rn, (unsigned)s, g);
}

// fp_div_s (src/fixed.wgsl:219-261): RNE(a * 2^s / b), s in 1..31, with the
// reference's saturation and flag rules.  Its 64-step restoring division
// yields the exact quotient and remainder of (|a| << s) / |b|; one 64-bit
// integer division gives the same pair.  Only the containment passes call it.
SN_HD int32_t fp_div_s(int32_t a, int32_t b, unsigned s, Dbg& g) {
  if (b == 0) {
    g.div0++;
    if (a == 0) return 0;
    return a > 0 ? I32_MAX : I32_MIN;
  }
  const bool neg = (a < 0) != (b < 0);
  const uint64_t ua = a < 0 ? (uint64_t)(0u - (uint32_t)a) : (uint64_t)(uint32_t)a;
  const uint64_t ub = b < 0 ? (uint64_t)(0u - (uint32_t)b) : (uint64_t)(uint32_t)b;
  const uint64_t n = ua << s;
  uint64_t q = n / ub;
  const uint64_t rem = n - q * ub, other = ub - rem;
  if (rem > other || (rem == other && (q & 1u))) q += 1;
  if (q > 0xffffffffull) { g.div++; return neg ? I32_MIN : I32_MAX; }
  const uint32_t q32 = (uint32_t)q;
  if (neg) {
    if (q32 > 2147483648u) { g.div++; return I32_MIN; }
    return (int32_t)(0u - q32);
  }
  if (q32 > 2147483647u) { g.div++; return I32_MAX; }
  return (int32_t)q32;
}

// ---- exact binary64 forms of fp_mul_s for bounded operands ----------------
// When |a * b| < 2^53, fp_mul_s(a, b, s) equals q_sat(q_rne((double)a *
// ((double)b * 2^-s))): the power-of-two scale is exact, the product of the
// two is the exact rational a*b / 2^s (its significand needs at most 53
// bits), and rint rounds it half to even (v_rndne_f64).  Saturation and the
// dbg[0] event then follow from the rounded value.  gfx950 issues binary64
// multiply, rint, min/max and the i32 <-> f64 conversions at the full VALU
// rate; the exact 64-bit integer product costs two 64-bit multiply-adds and a
// 64-bit add / shift / compare chain per multiply.  Every caller states the
// operand bound that makes its product exact.
SN_HD double q_rne(double x) { return __builtin_rint(x); }
// An integer-valued q with |q| < 2^52 saturated to i32; the event is the
// one fp_mul_s counts: q outside [I32_MIN, I32_MAX] <=> |q + 0.5| > 2^31
// (q + 0.5 is exact for |q| < 2^52).
SN_HD uint32_t q_ovf(double q) { return __builtin_fabs(q + 0.5) > 2147483648.0 ? 1u : 0u; }
SN_HD int32_t q_sat(double q, Dbg& g) {
  g.mul += q_ovf(q);
  return (int32_t)__builtin_fmin(__builtin_fmax(q, -2147483648.0), 2147483647.0);
}
// clampi(q_sat(q), lo, hi) for lo <= hi (clamping the saturated value or the
// unsaturated one gives the same result), with q_sat's event.
SN_HD int32_t q_sat_clamp(double q, int32_t lo, int32_t hi, Dbg& g) {
  g.mul += q_ovf(q);
  return (int32_t)__builtin_fmin(__builtin_fmax(q, (double)lo), (double)hi);
}
// The whole multiply in this form (for checks of the equivalence above).
SN_HD int32_t fp_mul_s_q(int32_t a, int32_t b, unsigned s, Dbg& g) {
  return q_sat(q_rne((double)a * __builtin_ldexp((double)b, -(int)s)), g);
}

// isqrt64 (src/fixed.wgsl:318-332) for a 32-bit input: floor(sqrt(n)).
SN_HD uint32_t isqrt32(uint32_t n) {
  uint32_t rem = 0, root = 0;
  for (int i = 15; i >= 0; --i) {
    rem = (rem << 2) | ((n >> (2 * i)) & 3u);
    uint32_t trial = (root << 2) + 1u;
    root <<= 1;
    if (rem >= trial) { rem -= trial; root |= 1u; }
  }
  return root;
}

// ---- MurmurHash3 pieces of the state hash (src/hash/statehash.wgsl:54-76) --
SN_HD uint32_t rotl32(uint32_t x, uint32_t r) { return (x << (r & 31u)) | (x >> ((32u - r) & 31u)); }
SN_HD uint32_t mm_step(uint32_t h, uint32_t k) {
  k *= 0xcc9e2d51u;
  k = rotl32(k, 15);
  k *= 0x1b873593u;
  h ^= k;
  h = rotl32(h, 13);
  return h * 5u + 0xe6546b64u;
}
SN_HD uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

}  // namespace sn

// ORACLE — test infrastructure only.  CPU restatement of the reference's
// Q16.16 integer core (src/fixed.wgsl).  Only tests/, the smoke check and the
// CPU-baseline leg of bench.py may load this code; the product never does.
//
// Every function states the reference lines it follows.  WGSL i32 `+ - *`
// wrap, so wrapping ops are spelled through uint32_t here.
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdint>

namespace oref {

constexpr int32_t I32_MAX = 2147483647;
constexpr int32_t I32_MIN = -2147483647 - 1;

// Debug event counters (src/fixed.wgsl:62-77).  The active world points this
// at its own 16 counters before running a pass.
extern std::atomic<uint32_t>* g_dbg;
inline void fp_flag(unsigned slot) { g_dbg[slot].fetch_add(1u, std::memory_order_relaxed); }

inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
inline int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
inline int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
inline int32_t clampi(int32_t x, int32_t lo, int32_t hi) { return x < lo ? lo : (x > hi ? hi : x); }
// WGSL clamp: min(max(x, lo), hi), which differs from clampi when lo > hi.
inline int32_t wclamp(int32_t x, int32_t lo, int32_t hi) { return std::min(std::max(x, lo), hi); }

// i64_shr_rne_sat (src/fixed.wgsl:146-164): floor shift by s in 1..31, round
// half to even on the exact remainder, saturate to i32 (dbg[0]) by the sign of v.
inline int32_t i64_shr_rne_sat(int64_t v, unsigned s) {
  int64_t q = v >> s;
  uint64_t rem = (uint64_t)v & ((1ull << s) - 1ull);
  uint64_t half = 1ull << (s - 1u);
  if (rem > half || (rem == half && (q & 1))) q += 1;
  if (q > (int64_t)I32_MAX || q < (int64_t)I32_MIN) {
    fp_flag(0);
    return v < 0 ? I32_MIN : I32_MAX;
  }
  return (int32_t)q;
}

// fp_mul_s (src/fixed.wgsl:168): exact 32x32->64 product, then the RNE shift.
inline int32_t fp_mul_s(int32_t a, int32_t b, unsigned s) {
  return i64_shr_rne_sat((int64_t)a * (int64_t)b, s);
}

// fp_add_sat / fp_neg_sat / fp_sub_sat / fp_abs (src/fixed.wgsl:175-200).
inline int32_t fp_add_sat(int32_t a, int32_t b) {
  int32_t s = wadd(a, b);
  if (((a ^ s) & (b ^ s)) < 0) { fp_flag(5); return a < 0 ? I32_MIN : I32_MAX; }
  return s;
}
inline int32_t fp_neg_sat(int32_t a) {
  if (a == I32_MIN) { fp_flag(5); return I32_MAX; }
  return -a;
}
inline int32_t fp_sub_sat(int32_t a, int32_t b) {
  int32_t s = wsub(a, b);
  if (((a ^ b) & (a ^ s)) < 0) { fp_flag(5); return a < 0 ? I32_MIN : I32_MAX; }
  return s;
}
inline int32_t fp_abs(int32_t a) { return a < 0 ? fp_neg_sat(a) : a; }

// fp_shr_rne (src/fixed.wgsl:205-213): RNE arithmetic shift of an i32, s in 1..31.
inline int32_t fp_shr_rne(int32_t x, unsigned s) {
  int32_t q = x >> s;
  uint32_t rem = (uint32_t)x & ((1u << s) - 1u);
  uint32_t half = 1u << (s - 1u);
  int32_t inc = 0;
  if (rem > half) inc = 1;
  else if (rem == half && (q & 1) == 1) inc = 1;
  return q + inc;
}

inline uint32_t u32abs(int32_t v) { return v < 0 ? (~(uint32_t)v) + 1u : (uint32_t)v; }

// fp_div_s (src/fixed.wgsl:219-261): correctly rounded RNE(a * 2^s / b) with the
// reference's flag and saturation rules.  The reference's 64-step restoring
// division computes the exact quotient Q = (|a| << s) / |b| and remainder; the
// same values come from 64-bit integer division here.
inline int32_t fp_div_s(int32_t a, int32_t b, unsigned s) {
  if (b == 0) {
    fp_flag(2);
    if (a == 0) return 0;
    return a > 0 ? I32_MAX : I32_MIN;
  }
  bool neg = (a < 0) != (b < 0);
  uint64_t ua = u32abs(a), ub = u32abs(b);
  uint64_t n = ua << s;
  uint64_t Q = n / ub, rem = n % ub;
  uint64_t other = ub - rem;
  bool incq = (rem > other) || (rem == other && (Q & 1ull));
  if (incq) Q += 1;
  if (Q > 0xFFFFFFFFull) { fp_flag(1); return neg ? I32_MIN : I32_MAX; }
  uint32_t q = (uint32_t)Q;
  if (neg) {
    if (q > 2147483648u) { fp_flag(1); return I32_MIN; }
    if (q == 2147483648u) return I32_MIN;
    return -(int32_t)q;
  }
  if (q > 2147483647u) { fp_flag(1); return I32_MAX; }
  return (int32_t)q;
}

inline int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }

// fp_divshift (src/fixed.wgsl:285-309): the solver divide.  clz-normalise den
// to [2^15, 2^16), Rn = 2^30 / dn (exact i32 divide), RNE(num * Rn >> (30+e-k)).
inline int32_t fp_divshift(int32_t num, int32_t den, unsigned k) {
  if (den == 0) {
    fp_flag(2);
    if (num == 0) return 0;
    return num > 0 ? I32_MAX : I32_MIN;
  }
  int32_t d = den, nn = num;
  if (d < 0) { fp_flag(1); d = fp_abs(d); nn = fp_neg_sat(nn); }
  int e = 16 - clz32((uint32_t)d);
  int32_t dn = e >= 0 ? (d >> e) : (d << (-e));
  int32_t Rn = 1073741824 / dn;
  int s = 30 + e - (int)k;
  if (s > 31) { fp_flag(1); return 0; }
  if (s < 1) {
    fp_flag(1);
    if (nn == 0) return 0;
    return nn > 0 ? I32_MAX : I32_MIN;
  }
  return i64_shr_rne_sat((int64_t)nn * (int64_t)Rn, (unsigned)s);
}

// isqrt64 (src/fixed.wgsl:318-332): floor(sqrt(n)) for an unsigned 64-bit n.
inline uint32_t isqrt64(uint64_t n) {
  uint64_t rem = 0;
  uint32_t root = 0;
  for (int i = 31; i >= 0; --i) {
    uint64_t bits = (n >> (2 * i)) & 3ull;
    rem = (rem << 2) | bits;
    uint64_t trial = ((uint64_t)root << 2) + 1ull;
    root <<= 1;
    if (rem >= trial) { rem -= trial; root |= 1u; }
  }
  return root;
}

}  // namespace oref

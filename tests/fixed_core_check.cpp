// CPU check of the Q16.16 core (tests only).
//
// Three things are compared on the same inputs:
//   * closed forms in exact __int128 arithmetic (what the reference's
//     functions are specified to compute: RNE shift with saturation, the
//     restoring divide's correctly rounded quotient, integer sqrt);
//   * the oracle's restatement (oracle/fixed_ref.hpp);
//   * the product's device header compiled for the host
//     (sunaengine_amd/csrc/fixed.hpp).
// Exhaustive over small operands and over every fp_divshift normalised
// denominator; randomised plus edge cases (I32_MIN, ties at half, s = 1, 31)
// elsewhere.  Prints "OK <checks>" or the first failure.
#include <cstdint>
#include <cstdio>
#include <random>

#include "../oracle/fixed_ref.hpp"
#include "../sunaengine_amd/csrc/fixed.hpp"

namespace oref {
std::atomic<uint32_t> dbgs[16];
std::atomic<uint32_t>* g_dbg = dbgs;
}  // namespace oref

static long checks = 0;
#define EXPECT(cond, ...)                 \
  do {                                    \
    ++checks;                             \
    if (!(cond)) {                        \
      std::printf("FAIL " __VA_ARGS__);   \
      std::printf("\n");                  \
      return 1;                           \
    }                                     \
  } while (0)

typedef __int128 i128;

static int32_t sat32(i128 v) {
  if (v > (i128)INT32_MAX) return INT32_MAX;
  if (v < (i128)INT32_MIN) return INT32_MIN;
  return (int32_t)v;
}
// RNE(v / 2^s) in exact arithmetic, saturated; the sign of v picks the rail
static int32_t ref_shr_rne_sat(i128 v, unsigned s) {
  i128 d = (i128)1 << s;
  i128 q = v >= 0 ? v / d : -((-v + d - 1) / d);  // floor
  i128 r = v - q * d;                             // [0, d)
  if (2 * r > d || (2 * r == d && (q & 1))) q += 1;
  if (q > INT32_MAX || q < INT32_MIN) return v < 0 ? INT32_MIN : INT32_MAX;
  return (int32_t)q;
}
static int32_t ref_div_s(int32_t a, int32_t b, unsigned s) {
  if (b == 0) return a == 0 ? 0 : (a > 0 ? INT32_MAX : INT32_MIN);
  i128 n = (i128)a * ((i128)1 << s);
  i128 bb = b;
  bool neg = (n < 0) != (bb < 0);
  i128 an = n < 0 ? -n : n, ab = bb < 0 ? -bb : bb;
  i128 q = an / ab, r = an % ab;
  if (2 * r > ab || (2 * r == ab && (q & 1))) q += 1;
  if (neg) q = -q;
  return sat32(q);
}

int main() {
  std::mt19937_64 rng(12345);
  const int32_t edge[] = {0, 1, -1, 2, -2, 3, 65535, 65536, -65536, 32767, 32768, -32768,
                          INT32_MAX, INT32_MIN, INT32_MAX - 1, INT32_MIN + 1, 1 << 20,
                          -(1 << 20), 1 << 30, -(1 << 30), 163840, -163840};
  auto rnd = [&]() -> int32_t {
    switch (rng() % 4) {
      case 0: return (int32_t)(rng() & 0xffff) - 32768;
      case 1: return (int32_t)(rng() & 0xffffff) - (1 << 23);
      case 2: return (int32_t)rng();
      default: return edge[rng() % (sizeof(edge) / sizeof(edge[0]))];
    }
  };
  sn::Dbg g;
  // ---- fp_mul_s
  for (int it = 0; it < 400000; ++it) {
    int32_t a = rnd(), b = rnd();
    unsigned s = 1 + (unsigned)(rng() % 31);
    int32_t want = ref_shr_rne_sat((i128)a * b, s);
    EXPECT(oref::fp_mul_s(a, b, s) == want, "oracle fp_mul_s(%d,%d,%u)", a, b, s);
    EXPECT(sn::fp_mul_s(a, b, s, g) == want, "device fp_mul_s(%d,%d,%u)", a, b, s);
  }
  // the saturation count of fp_mul_s (dbg[0]) matches the exact quotient
  for (int it = 0; it < 200000; ++it) {
    int32_t a = rnd(), b = rnd();
    unsigned s = 1 + (unsigned)(rng() % 31);
    i128 v = (i128)a * b;
    i128 q = v >> s;                                         // floor
    i128 rem = v - (q << s), half = (i128)1 << (s - 1);
    if (rem > half || (rem == half && (q & 1))) q += 1;     // round half to even
    bool sat = q > INT32_MAX || q < INT32_MIN;
    sn::Dbg g1;
    sn::fp_mul_s(a, b, s, g1);
    EXPECT(g1.mul == (sat ? 1u : 0u), "saturation count fp_mul_s(%d,%d,%u)", a, b, s);
  }
  // the binary64 form: equal value and equal saturation event whenever
  // |a * b| < 2^53 (b drawn below 2^(53 - bits(a))), every shift, with ties
  // and both rails; and the fused clamp equals clampi of the saturated value
  for (int it = 0; it < 600000; ++it) {
    int32_t a = rnd();
    int32_t b = rnd();
    const int abits = a == 0 ? 0 : 64 - __builtin_clzll((uint64_t)(a < 0 ? -(int64_t)a : (int64_t)a));
    const int room = 53 - abits;   // |b| < 2^room keeps |a b| < 2^53
    if (room < 31) b = (int32_t)((int64_t)b >> (31 - room));
    if (it % 5 == 0) {   // an exact tie at half of the last kept bit
      const unsigned s0 = 1 + (unsigned)(rng() % 31);
      a = (int32_t)((((int64_t)(rng() % 4001) - 2000) << s0) + ((int64_t)1 << (s0 - 1)));
      if ((int64_t)a != (((int64_t)(a) >> s0) << s0) + ((int64_t)1 << (s0 - 1))) continue;
      b = 1;
    }
    const i128 v = (i128)a * b;
    if (v >= ((i128)1 << 53) || v <= -((i128)1 << 53)) continue;
    const unsigned s = 1 + (unsigned)(rng() % 31);
    sn::Dbg g1, g2;
    const int32_t want = sn::fp_mul_s(a, b, s, g1);
    EXPECT(sn::fp_mul_s_q(a, b, s, g2) == want && g1.mul == g2.mul, "binary64 fp_mul_s(%d,%d,%u)", a, b, s);
    const int32_t lo = -(int32_t)(rng() % 0x7fffffffu), hi = (int32_t)(rng() % 0x7fffffffu);
    sn::Dbg g3;
    const double q = sn::q_rne((double)a * __builtin_ldexp((double)b, -(int)s));
    EXPECT(sn::q_sat_clamp(q, lo, hi, g3) == sn::clampi(want, lo, hi) && g3.mul == g1.mul,
           "binary64 clamp(%d,%d,%u,[%d,%d])", a, b, s, lo, hi);
  }
  for (unsigned s = 1; s <= 31; ++s)
    for (int32_t a : {INT32_MIN, INT32_MAX, INT32_MIN + 1, -1, 1, 0})
      for (int32_t b : {-(1 << 21), (1 << 21) - 1, 1 << 20, -3, 3, 0}) {
        sn::Dbg g1, g2;
        EXPECT(sn::fp_mul_s_q(a, b, s, g2) == sn::fp_mul_s(a, b, s, g1) && g1.mul == g2.mul,
               "binary64 rails(%d,%d,%u)", a, b, s);
      }
  // exhaustive small operands, all shifts
  for (int32_t a = -300; a <= 300; ++a)
    for (int32_t b = -40; b <= 40; ++b)
      for (unsigned s = 1; s <= 31; s += 3) {
        int32_t want = ref_shr_rne_sat((i128)a * b, s);
        EXPECT(oref::fp_mul_s(a, b, s) == want && sn::fp_mul_s(a, b, s, g) == want, "small mul");
      }
  // ties at exactly half, both signs
  for (unsigned s = 1; s <= 31; ++s)
    for (int k = -50; k <= 50; ++k) {
      int64_t v = ((int64_t)k << s) + ((int64_t)1 << (s - 1));
      if (v > INT32_MAX || v < INT32_MIN) continue;
      int32_t want = ref_shr_rne_sat(v, s);
      EXPECT(oref::fp_mul_s((int32_t)v, 1, s) == want && sn::fp_mul_s((int32_t)v, 1, s, g) == want, "tie");
      EXPECT(sn::fp_shr_rne((int32_t)v, s) == want && oref::fp_shr_rne((int32_t)v, s) == want, "shr tie");
    }
  // ---- fp_shr_rne over random i32
  for (int it = 0; it < 200000; ++it) {
    int32_t x = rnd();
    unsigned s = 1 + (unsigned)(rng() % 31);
    int32_t want = ref_shr_rne_sat((i128)x, s);
    EXPECT(oref::fp_shr_rne(x, s) == want && sn::fp_shr_rne(x, s) == want, "shr %d %u", x, s);
  }
  // ---- saturating add/sub/neg/abs
  for (int it = 0; it < 200000; ++it) {
    int32_t a = rnd(), b = rnd();
    EXPECT(oref::fp_add_sat(a, b) == sat32((i128)a + b) && sn::fp_add_sat(a, b, g) == sat32((i128)a + b), "add");
    EXPECT(oref::fp_sub_sat(a, b) == sat32((i128)a - b) && sn::fp_sub_sat(a, b, g) == sat32((i128)a - b), "sub");
    EXPECT(oref::fp_neg_sat(a) == sat32(-(i128)a) && sn::fp_neg_sat(a, g) == sat32(-(i128)a), "neg");
    i128 ab = a < 0 ? -(i128)a : (i128)a;
    EXPECT(oref::fp_abs(a) == sat32(ab) && sn::fp_abs(a, g) == sat32(ab), "abs");
  }
  // ---- fp_div_s: oracle and product against the exact quotient
  for (int it = 0; it < 200000; ++it) {
    int32_t a = rnd(), b = rnd();
    unsigned s = (unsigned)(rng() % 32);
    if (s == 0) s = 16;
    if (it % 4 == 1) b = (int32_t)(rng() % 9000) - 4500;  // small divisors: saturation and ties
    if (it % 4 == 2) a = (int32_t)(rng() % 9000) - 4500;
    const int32_t want = ref_div_s(a, b, s);
    EXPECT(oref::fp_div_s(a, b, s) == want, "div_s(%d,%d,%u)", a, b, s);
    EXPECT(sn::fp_div_s(a, b, s, g) == want, "sn div_s(%d,%d,%u)", a, b, s);
  }
  for (int32_t a = -600; a <= 600; ++a)
    for (int32_t b = -40; b <= 40; ++b)
      EXPECT(sn::fp_div_s(a, b, 16, g) == ref_div_s(a, b, 16) && sn::fp_div_s(a, b, 1, g) == ref_div_s(a, b, 1),
             "sn div_s small(%d,%d)", a, b);
  EXPECT(sn::fp_div_s(INT32_MIN, 1, 1, g) == INT32_MIN && sn::fp_div_s(INT32_MIN, -1, 1, g) == INT32_MAX &&
             sn::fp_div_s(-65536, 65536, 31, g) == INT32_MIN && sn::fp_div_s(65536, 65536, 31, g) == INT32_MAX,
         "div_s rails");
  // ---- fp_divshift: product == oracle everywhere; closed-form structure
  for (int32_t dn = 32768; dn < 65536; ++dn) {
    int32_t want = (int32_t)(1073741824LL / dn);
    EXPECT(sn::recip30(dn) == want, "recip30(%d)", dn);
  }
  for (int it = 0; it < 400000; ++it) {
    int32_t num = rnd(), den = rnd();
    unsigned k = 1 + (unsigned)(rng() % 26);
    int32_t a = oref::fp_divshift(num, den, k);
    int32_t b = sn::fp_divshift(num, den, k, g);
    EXPECT(a == b, "divshift(%d,%d,%u) oracle %d device %d", num, den, k, a, b);
    // closed form of the normalised reciprocal path
    if (den >= 32768) {
      int e = 16 - __builtin_clz((uint32_t)den);
      int32_t dn = e >= 0 ? den >> e : den << -e;
      int32_t Rn = (int32_t)(1073741824LL / dn);
      int sh = 30 + e - (int)k;
      if (sh >= 1 && sh <= 31) {
        EXPECT(a == ref_shr_rne_sat((i128)num * Rn, (unsigned)sh), "divshift closed form");
        // relative error bound of the 15-bit reciprocal, for den >= 2^15
        long double exact = (long double)num * (long double)(1ULL << k) / (long double)den;
        // |error| <= |exact| * 2^-14 (reciprocal and normalisation) + 0.5 (final RNE)
        if (a != INT32_MAX && a != INT32_MIN) {
          long double err = a - exact, mag = exact < 0 ? -exact : exact;
          if (err < 0) err = -err;
          EXPECT(err <= mag / 16384.0L + 0.5L + 1e-9L, "divshift error num=%d den=%d k=%u got=%d exact=%Lf", num,
                 den, k, a, exact);
        }
      }
    }
  }
  // sign symmetry (src/fixed.wgsl:265-276)
  for (int it = 0; it < 100000; ++it) {
    int32_t num = rnd(), den = (int32_t)(32768 + rng() % 2000000000u);
    if (num == INT32_MIN) continue;
    EXPECT(sn::fp_divshift(-num, den, 16, g) == -sn::fp_divshift(num, den, 16, g) ||
               sn::fp_divshift(num, den, 16, g) == INT32_MIN,
           "divshift symmetry");
  }
  // ---- isqrt
  for (uint64_t n = 0; n < 300000; ++n) {
    uint32_t r = oref::isqrt64(n);
    EXPECT((uint64_t)r * r <= n && (uint64_t)(r + 1) * (r + 1) > n, "isqrt64(%llu)", (unsigned long long)n);
    EXPECT(sn::isqrt32((uint32_t)n) == r, "isqrt32");
  }
  for (int it = 0; it < 200000; ++it) {
    uint32_t n = (uint32_t)rng();
    uint32_t r = oref::isqrt64(n);
    EXPECT((uint64_t)r * r <= n && (uint64_t)(r + 1) * (r + 1) > n, "isqrt64 big");
    EXPECT(sn::isqrt32(n) == r, "isqrt32 big");
  }
  std::printf("OK %ld\n", checks);
  return 0;
}

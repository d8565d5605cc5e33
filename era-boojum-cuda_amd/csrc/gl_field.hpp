// Goldilocks field arithmetic, p = 2^64 - 2^32 + 1, for host and gfx950.
//
// Semantics follow the reference's base field (native/goldilocks.cuh:90-185),
// with one difference in representation: the reference lets any u64 stand for
// its residue, while here every value inside a transform is CANONICAL, in
// [0, p).  Inputs are canonicalized once when loaded; add/sub require and
// return canonical values (which removes the double-wrap corrections a
// non-canonical add needs), mul and mul_pow2 accept any u64 and return
// canonical values.  The implementation is derived from the two identities
//   2^64 = 2^32 - 1 (mod p)      2^96 = -1 (mod p)
// and is written for AMD's 32-bit VALU: 64x64 products are four
// v_mad_u64_u32 steps, carries are read back from the add/sub carry-out.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GL_FN __host__ __device__ __forceinline__
#else
#define GL_FN inline
#endif

namespace gl {

constexpr uint64_t P = 0xFFFFFFFF00000001ull;
constexpr uint64_t EPS = 0x00000000FFFFFFFFull;  // 2^64 mod p

// Canonical representative in [0, p) of any u64.  x >= p exactly when
// x + EPS wraps, so no comparison with p is needed.
GL_FN uint64_t canon(uint64_t x) {
  const uint64_t t = x + EPS;
  return t < x ? t : x;
}

// a + b mod p for CANONICAL a, b; result canonical.  b + EPS = b - p + 2^64
// does not wrap (b <= p - 1 = 2^64 - EPS - 1), so r = a + (b + EPS) wraps
// (r < a) exactly when a + b >= p, and r is then a + b - p: one 64-bit
// compare decides, with no second wrap test.
GL_FN uint64_t add(uint64_t a, uint64_t b) {
  const uint64_t s = a + b, r = a + (b + EPS);
  return r < a ? r : s;
}

// a - b mod p for CANONICAL a, b; result canonical.  The borrow comes from
// the 32-bit subtract chain itself (no separate comparison); on a borrow
// a - b + 2^64 >= 2^64 - p + 1, so adding p (mod 2^64) lands in [1, p).
GL_FN uint64_t sub(uint64_t a, uint64_t b) {
  unsigned b0, b1;
  const uint32_t lo = __builtin_subc((uint32_t)a, (uint32_t)b, 0u, &b0);
  const uint32_t hi = __builtin_subc((uint32_t)(a >> 32), (uint32_t)(b >> 32), b0, &b1);
  const uint64_t d = ((uint64_t)hi << 32) | lo;
  return b1 ? d + P : d;
}

GL_FN uint64_t neg(uint64_t a) { return sub(0ull, a); }

// Canonical value of x + d * 2^64 for d in {-1, 0, +1} (minus / plus flags),
// given bounds that make x + EPS (plus) and x + P (minus) canonical.
GL_FN uint64_t fold(uint64_t x, bool plus, bool minus) {
  const bool add_eps = plus | (!minus & (x >= P));
  const uint32_t lo = add_eps ? 0xFFFFFFFFu : (minus ? 1u : 0u);
  const uint32_t hi = minus ? 0xFFFFFFFFu : 0u;
  return x + (((uint64_t)hi << 32) | lo);
}

// Reduce l0 + l1 2^32 + l2 2^64 + l3 2^96 (32-bit limbs) to canonical:
// 2^64 = 2^32 - 1 and 2^96 = -1 give (l1 + l2) 2^32 + l0 - (l2 + l3), a value
// in (-2^33, 2^65) that one fold corrects.
GL_FN uint64_t reduce_limbs(uint32_t l0, uint32_t l1, uint32_t l2, uint32_t l3) {
  unsigned c, sc, b0, b1;
  uint32_t hi = __builtin_addc(l1, l2, 0u, &c);
  const uint32_t ssum = __builtin_addc(l2, l3, 0u, &sc);
  const uint32_t lo = __builtin_subc(l0, ssum, 0u, &b0);
  hi = __builtin_subc(hi, sc, b0, &b1);
  // c set, b1 clear: (hi:lo) <= 2^64 - 2^33, so + EPS stays below p;
  // b1 set, c clear: the value is above -2^33, so + p is canonical.
  return fold(((uint64_t)hi << 32) | lo, c && !b1, b1 && !c);
}

// a * b mod p for any u64 a, b; result canonical.  The 128-bit product is
// accumulated column by column through the 64-bit addend of four
// 32x32+64 multiply-adds; every partial sum stays below 2^64
// ((2^32 - 1)^2 + 2 (2^32 - 1) = 2^64 - 1), so no carry needs checking:
//   t1 = a0 b1 + hi(a0 b0),  t2 = a1 b0 + lo(t1),  t3 = a1 b1 + hi(t1),
//   t4 = t3 + hi(t2);  limbs lo(a0 b0), lo(t2), lo(t4), hi(t4).
GL_FN uint64_t mul(uint64_t a, uint64_t b) {
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
  const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  const uint64_t t0 = (uint64_t)a0 * b0;
  const uint64_t t1 = (uint64_t)a0 * b1 + (t0 >> 32);
  const uint64_t t2 = (uint64_t)a1 * b0 + (uint32_t)t1;
  const uint64_t t3 = (uint64_t)a1 * b1 + (t1 >> 32);
  const uint64_t t4 = t3 + (t2 >> 32);
  return reduce_limbs((uint32_t)t0, (uint32_t)t2, (uint32_t)t4, (uint32_t)(t4 >> 32));
}

// Sum of products kept unreduced in five 32-bit limbs,
// l0 + l1 2^32 + l2 2^64 + l3 2^96 + l4 2^128: each product costs the four
// multiply-adds of mul plus a 5-step carry chain, and the reduction is paid
// once per sum.  l4 counts carries out of 128 bits, so it stays below the
// number of products (at most 2^32 - 2 of them).
struct Acc {
  uint32_t l0 = 0, l1 = 0, l2 = 0, l3 = 0, l4 = 0;
};

// acc += a * b for any u64 a, b (the product columns as in mul).
GL_FN void mac(Acc &acc, uint64_t a, uint64_t b) {
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
  const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  const uint64_t t0 = (uint64_t)a0 * b0;
  const uint64_t t1 = (uint64_t)a0 * b1 + (t0 >> 32);
  const uint64_t t2 = (uint64_t)a1 * b0 + (uint32_t)t1;
  const uint64_t t3 = (uint64_t)a1 * b1 + (t1 >> 32);
  const uint64_t t4 = t3 + (t2 >> 32);
  unsigned c;
  acc.l0 = __builtin_addc(acc.l0, (uint32_t)t0, 0u, &c);
  acc.l1 = __builtin_addc(acc.l1, (uint32_t)t2, c, &c);
  acc.l2 = __builtin_addc(acc.l2, (uint32_t)t4, c, &c);
  acc.l3 = __builtin_addc(acc.l3, (uint32_t)(t4 >> 32), c, &c);
  acc.l4 += c;
}

// Canonical value of the sum: 2^128 = 2^96 2^32 = -2^32 (mod p), and
// l4 2^32 < p for l4 < 2^32 - 1.
GL_FN uint64_t acc_reduce(const Acc &acc) {
  return sub(reduce_limbs(acc.l0, acc.l1, acc.l2, acc.l3), (uint64_t)acc.l4 << 32);
}

// x * 2^E for 0 < E < 64, canonical result, any u64 x.
template <unsigned E> GL_FN uint64_t pow2_direct(uint64_t x) {
  static_assert(E > 0 && E < 64, "");
  const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32);
  if constexpr (E < 32) {
    // x << E = y0 + y1 2^32 + y2 2^64 = (y1 + y2) 2^32 + y0 - y2, y2 < 2^E
    const uint32_t y0 = x0 << E;
    const uint32_t y1 = (x1 << E) | (x0 >> (32 - E));
    const uint32_t y2 = x1 >> (32 - E);
    unsigned c, b0, b1;
    uint32_t hi = __builtin_addc(y1, y2, 0u, &c);
    const uint32_t lo = __builtin_subc(y0, y2, 0u, &b0);
    hi = __builtin_subc(hi, 0u, b0, &b1);
    // value = c 2^64 + (hi:lo) - b1 2^64; with c set, hi < y2 < 2^E keeps
    // (hi:lo) + EPS below p; c = 0, b1 = 1 cannot happen
    return fold(((uint64_t)hi << 32) | lo, c && !b1, false);
  } else if constexpr (E == 32) {
    return reduce_limbs(0u, x0, x1, 0u);  // x 2^32 = x0 2^32 + x1 2^64
  } else {
    // 2^E = -2^-K with K = 96 - E in (32, 64): split x = q 2^K + r, then
    // x 2^E = r 2^E - q with r 2^E < 2^96:  r 2^E = (Lhi 2^32) + H 2^64,
    // H 2^64 = H 2^32 - H, so x 2^E = (Lhi + H) 2^32 - (H + q).
    constexpr unsigned K = 96 - E;
    const uint32_t lhi = x0 << (E - 32);
    const uint32_t rhi = x1 & ((1u << (K - 32)) - 1u);
    const uint32_t h = (uint32_t)((((uint64_t)rhi << 32) | x0) >> (K - 32));
    const uint32_t q = x1 >> (K - 32);
    unsigned c, sc, b0, b1;
    uint32_t hi = __builtin_addc(lhi, h, 0u, &c);
    const uint32_t ssum = __builtin_addc(h, q, 0u, &sc);
    const uint32_t lo = __builtin_subc(0u, ssum, 0u, &b0);
    hi = __builtin_subc(hi, sc, b0, &b1);
    return fold(((uint64_t)hi << 32) | lo, c && !b1, b1 && !c);
  }
}

// x * 2^-K for 0 < K <= 32, canonical result, any u64 x:
// x = q 2^K + r, r 2^-K = -r 2^(96-K) = -(r << (32-K)) (2^32 - 1), so
// x 2^-K = q + r' - r' 2^32 with r' = r << (32 - K) < 2^32.
template <unsigned K> GL_FN uint64_t pow2_inverse(uint64_t x) {
  static_assert(K > 0 && K <= 32, "");
  const uint32_t rp = K == 32 ? (uint32_t)x : ((uint32_t)x << ((32 - K) % 32));
  const uint64_t a = (x >> K) + rp;  // < 2^(64-K) + 2^32, no wrap
  const uint32_t ahi = (uint32_t)(a >> 32);
  unsigned b;
  const uint32_t hi = __builtin_subc(ahi, rp, 0u, &b);
  const uint64_t v = ((uint64_t)hi << 32) | (uint32_t)a;
  // no borrow: v < 2^(64-K) + 2^32 < p; borrow: v - 2^64 in (-2^64, 0), + p
  return b ? v + P : v;
}

// x * 2^e mod p for a compile-time 0 <= e < 192, as sign * value with value
// canonical (2^96 = -1 and 2^e = 2^(e-192)); callers fold the sign into the
// surrounding add/sub.  e = 0 passes x through and expects a canonical x.
template <unsigned e> struct Pow2 {
  static constexpr unsigned E = e % 192u;
  static constexpr bool NEG = (E >= 64 && E < 96) || (E >= 96 && E < 160);
  GL_FN static uint64_t value(uint64_t x) {
    if constexpr (E == 0) return x;
    else if constexpr (E < 64) return pow2_direct<E>(x);
    else if constexpr (E < 96) return pow2_inverse<96 - E>(x);
    else if constexpr (E == 96) return x;  // -x, sign in NEG
    else if constexpr (E < 160) return pow2_direct<E - 96>(x);
    else return pow2_inverse<192 - E>(x);
  }
};

// x * 2^E mod p, canonical, for 0 <= E < 192 (canonical x when E % 96 == 0).
template <unsigned E> GL_FN uint64_t mul_pow2(uint64_t x) {
  const uint64_t v = Pow2<E>::value(x);
  return Pow2<E>::NEG ? neg(v) : v;
}

GL_FN uint64_t pow(uint64_t b, uint64_t e) {
  uint64_t r = 1;
  while (e) {
    if (e & 1) r = mul(r, b);
    b = mul(b, b);
    e >>= 1;
  }
  return canon(r);
}

// x^(2^k) by k squarings.
GL_FN uint64_t sqn(uint64_t x, unsigned k) {
  for (unsigned i = 0; i < k; ++i) x = mul(x, x);
  return x;
}

// a^(p - 2) = 1 / a (0 for a = 0) by an addition chain: with
// t_k = a^(2^k - 1), t_{j+k} = t_j^(2^k) t_k, so t_31 costs 30 squarings and
// 6 products; then p - 2 = 2^64 - 2^32 - 1 = (2^31 - 1) 2^33 + (2^32 - 1)
// gives a^(p-2) = t_31^(2^33) t_32 with t_32 = t_31^2 a: 64 squarings and
// 8 products in all, against 63 + 63 for square-and-multiply.
GL_FN uint64_t inv(uint64_t a) {
  const uint64_t t1 = canon(a);
  const uint64_t t2 = mul(mul(t1, t1), t1);
  const uint64_t t3 = mul(mul(t2, t2), t1);
  const uint64_t t6 = mul(sqn(t3, 3), t3);
  const uint64_t t12 = mul(sqn(t6, 6), t6);
  const uint64_t t24 = mul(sqn(t12, 12), t12);
  const uint64_t t30 = mul(sqn(t24, 6), t6);
  const uint64_t t31 = mul(mul(t30, t30), t1);
  const uint64_t t32 = mul(mul(t31, t31), t1);
  return mul(sqn(t31, 33), t32);
}

// Constants of the boojum convention.  g = 7 generates F_p^*; the radix-2
// subgroup generator is 7^((p-1)/2^32), and omega_{2^k} is its 2^(32-k)-th
// power, so every omega_{2^k} with k <= 6 is a power of two:
// omega_64 = 2^39 (forward) and omega_64^-1 = 2^153.
constexpr uint64_t GENERATOR = 7;
constexpr uint64_t OMEGA_2_32 = 0x185629DCDA58878Cull;
constexpr unsigned OMEGA_64_LOG2 = 39;

// Exponent e in [0, 192) with omega_{2^a}^t = 2^e (forward) or its inverse.
constexpr unsigned root_pow2_exp(unsigned a, unsigned t, bool inverse) {
  // omega_{2^a} = omega_64^(2^(6-a)) = 2^(39 * 2^(6-a)).
  unsigned base = (OMEGA_64_LOG2 << (6 - a)) % 192u;
  unsigned e = (base * (t % (1u << a))) % 192u;
  return inverse ? (192u - e) % 192u : e;
}

}  // namespace gl

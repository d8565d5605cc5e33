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
// and is written for AMD's 32-bit VALU: products are v_mad_u64_u32 chains
// (whose 64-bit addend also applies 2^64 = 2^32 - 1), and carries are read
// back from the add/sub and multiply-add carry-outs.
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

#if defined(__HIP_DEVICE_COMPILE__)
// gfx950 building blocks whose carries stay in lane masks (SGPR pairs).  A
// VALU instruction that reads a lane mask written by another VALU
// instruction needs two wait states; every block below that consumes such a
// mask as a carry-in starts with s_nop 1, and masks read by SALU first need
// none.  The blocks hold no memory operations.
namespace dev {

// a * b + c as 65 bits: the low 64 bits, and bit 64 in `carry`
// (v_mad_u64_u32 writes the carry out of its 64-bit addition to an SGPR pair).
__device__ __forceinline__ uint64_t mad_carry(uint32_t a, uint32_t b, uint64_t c, uint64_t &carry) {
  uint64_t r;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(carry) : "v"(a), "v"(b), "v"(c));
  return r;
}

// mad_carry with b in an SGPR (b the same in every lane).
__device__ __forceinline__ uint64_t mad_carry_s(uint32_t a, uint32_t b, uint64_t c, uint64_t &carry) {
  uint64_t r;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(carry) : "v"(a), "s"(b), "v"(c));
  return r;
}

// z - h - cin (cin a lane mask) with the final borrow in `borrow`.
__device__ __forceinline__ uint64_t sub32_borrow(uint64_t z, uint32_t h, uint64_t cin, uint64_t &borrow) {
  uint32_t lo, hi;
  uint64_t b0;
  asm("s_nop 1\n\t"
      "v_subb_co_u32_e64 %[lo], %[b0], %[zlo], %[h], %[cin]\n\t"
      "s_nop 1\n\t"
      "v_subb_co_u32_e64 %[hi], %[bo], %[zhi], 0, %[b0]"
      : [lo] "=&v"(lo), [hi] "=&v"(hi), [b0] "=&s"(b0), [bo] "=&s"(borrow)
      : [zlo] "v"((uint32_t)z), [zhi] "v"((uint32_t)(z >> 32)), [h] "v"(h), [cin] "s"(cin));
  return ((uint64_t)hi << 32) | lo;
}

// z - h with the borrow in `borrow`.
__device__ __forceinline__ uint64_t sub32(uint64_t z, uint32_t h, uint64_t &borrow) {
  uint32_t lo, hi;
  uint64_t b0;
  asm("v_sub_co_u32_e64 %[lo], %[b0], %[zlo], %[h]\n\t"
      "s_nop 1\n\t"
      "v_subb_co_u32_e64 %[hi], %[bo], %[zhi], 0, %[b0]"
      : [lo] "=&v"(lo), [hi] "=&v"(hi), [b0] "=&s"(b0), [bo] "=&s"(borrow)
      : [zlo] "v"((uint32_t)z), [zhi] "v"((uint32_t)(z >> 32)), [h] "v"(h));
  return ((uint64_t)hi << 32) | lo;
}

// Canonical value of x + (plus - minus) 2^64 for lane masks plus, minus,
// given the bounds of fold() below: a set plus alone adds EPS, a set minus
// alone adds P (subtracts EPS), and otherwise x >= P adds EPS.
__device__ __forceinline__ uint64_t fold_masks(uint64_t x, uint64_t plus, uint64_t minus) {
  uint32_t lo, hi;
  uint64_t ge, pl, mi;
  asm("v_cmp_le_u64_e64 %[ge], %[P], %[x]\n\t"
      "s_andn2_b64 %[pl], %[cp], %[cm]\n\t"
      "s_andn2_b64 %[mi], %[cm], %[cp]\n\t"
      "s_andn2_b64 %[ge], %[ge], %[mi]\n\t"
      "s_or_b64 %[ge], %[ge], %[pl]\n\t"
      "v_cndmask_b32_e64 %[lo], 0, 1, %[mi]\n\t"
      "v_cndmask_b32_e64 %[hi], 0, -1, %[mi]\n\t"
      "v_cndmask_b32_e64 %[lo], %[lo], -1, %[ge]"
      : [lo] "=&v"(lo), [hi] "=&v"(hi), [ge] "=&s"(ge), [pl] "=&s"(pl), [mi] "=&s"(mi)
      : [x] "v"(x), [P] "s"(P), [cp] "s"(plus), [cm] "s"(minus)
      : "scc");
  return x + (((uint64_t)hi << 32) | lo);
}

// Canonical value of x + plus 2^64 (plus a lane mask), given x + EPS < p
// wherever plus is set.
__device__ __forceinline__ uint64_t fold_plus(uint64_t x, uint64_t plus) {
  uint32_t lo;
  uint64_t ge;
  asm("v_cmp_le_u64_e64 %[ge], %[P], %[x]\n\t"
      "s_or_b64 %[ge], %[ge], %[cp]\n\t"
      "v_cndmask_b32_e64 %[lo], 0, -1, %[ge]"
      : [lo] "=&v"(lo), [ge] "=&s"(ge)
      : [x] "v"(x), [P] "s"(P), [cp] "s"(plus)
      : "scc");
  return x + (uint64_t)lo;
}

}  // namespace dev
#endif

// a * b mod p for any u64 a, b; result canonical.  Five 32x32+64
// multiply-adds, the last one applying 2^64 = EPS:
//   t0 = a0 b0,  t1 = a0 b1 + hi(t0)            (< 2^64)
//   t2 = a1 b0 + t1 = T2 + c 2^64               (c: carry out, 65-bit sum)
//   t3 = a1 b1 + hi(T2)                         (< 2^64)
// so a b = X + lo(t3) 2^64 + (hi(t3) + c) 2^96 with X = lo(t0) + lo(T2) 2^32,
// and with 2^64 = EPS, 2^96 = -1:
//   a b = X + lo(t3) EPS - hi(t3) - c = Z + c2 2^64 - hi(t3) - c,
// where Z + c2 2^64 = X + lo(t3) EPS is the fifth multiply-add.  Then
// W = Z - hi(t3) - c with final borrow bw, and a b = W + (c2 - bw) 2^64:
// c2 alone leaves W <= 2^64 - 2^33 (W + EPS < p); bw alone leaves
// W >= 2^64 - 2^32 (W - EPS canonical); both or neither leave W for the
// canonical check.  On gfx950 the carries stay in lane masks.
GL_FN uint64_t mul(uint64_t a, uint64_t b) {
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
  const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  const uint64_t t0 = (uint64_t)a0 * b0;
  const uint64_t t1 = (uint64_t)a0 * b1 + (t0 >> 32);
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t c, c2, bw;
  const uint64_t t2 = dev::mad_carry(a1, b0, t1, c);
  const uint64_t t3 = (uint64_t)a1 * b1 + (t2 >> 32);
  const uint64_t z = dev::mad_carry((uint32_t)t3, 0xFFFFFFFFu, (t2 << 32) | (uint32_t)t0, c2);
  const uint64_t w = dev::sub32_borrow(z, (uint32_t)(t3 >> 32), c, bw);
  return dev::fold_masks(w, c2, bw);
#else
  const uint64_t p2 = (uint64_t)a1 * b0;
  const uint64_t t2 = p2 + t1;
  const bool c = t2 < p2;
  const uint64_t t3 = (uint64_t)a1 * b1 + (t2 >> 32);
  const uint64_t x = (t2 << 32) | (uint32_t)t0;
  const uint64_t z = (uint64_t)(uint32_t)t3 * EPS + x;
  const bool c2 = z < x;
  unsigned b0w, bw;
  const uint32_t lo = __builtin_subc((uint32_t)z, (uint32_t)(t3 >> 32), (unsigned)c, &b0w);
  const uint32_t hi = __builtin_subc((uint32_t)(z >> 32), 0u, b0w, &bw);
  return fold(((uint64_t)hi << 32) | lo, c2 && !bw, bw && !c2);
#endif
}

// mul for a b that is the same in every lane of the wave (the caller's
// promise: lanes of another b get lane 0's).  On gfx950 b is taken into SGPRs,
// so a loop-invariant b hoisted out of a loop costs no VGPRs.
GL_FN uint64_t mul_uniform(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
  const uint32_t b0 = __builtin_amdgcn_readfirstlane((uint32_t)b);
  const uint32_t b1 = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
  const uint64_t t0 = (uint64_t)a0 * b0;
  const uint64_t t1 = (uint64_t)a0 * b1 + (t0 >> 32);
  uint64_t c, c2, bw;
  const uint64_t t2 = dev::mad_carry_s(a1, b0, t1, c);
  const uint64_t t3 = (uint64_t)a1 * b1 + (t2 >> 32);
  const uint64_t z = dev::mad_carry((uint32_t)t3, 0xFFFFFFFFu, (t2 << 32) | (uint32_t)t0, c2);
  const uint64_t w = dev::sub32_borrow(z, (uint32_t)(t3 >> 32), c, bw);
  return dev::fold_masks(w, c2, bw);
#else
  return mul(a, b);
#endif
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
//
// Device form (gfx950), one multiply-add applying 2^64 = EPS:
//   E < 32:  x 2^E = y + h 2^64, y = x << E (mod 2^64), h = hi(x) >> (32 - E),
//            so x 2^E = y + h EPS = Z + c 2^64 (65-bit multiply-add), and
//            Z + EPS < p where c is set (h < 2^31 gives Z < 2^63).
//   E >= 32: with e = E - 32, x 2^E = L + H 2^64, L = lo(x) 2^E (mod 2^64),
//            H = x >> (32 - e) < 2^E; H 2^64 = lo(H) EPS - hi(H), so
//            x 2^E = L + lo(H) EPS - hi(H) = Z + c 2^64 - hi(H), folded as in mul.
template <unsigned E> GL_FN uint64_t pow2_direct(uint64_t x) {
  static_assert(E > 0 && E < 64, "");
  const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32);
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t c;
  if constexpr (E < 32) {
    const uint64_t z = dev::mad_carry(x1 >> (32 - E), 0xFFFFFFFFu, x << E, c);
    return dev::fold_plus(z, c);
  } else {
    constexpr unsigned e = E - 32;
    const uint64_t lo_part = (uint64_t)(x0 << e) << 32;
    if constexpr (e == 0) {
      return dev::fold_plus(dev::mad_carry(x1, 0xFFFFFFFFu, lo_part, c), c);
    } else {
      const uint64_t z = dev::mad_carry((uint32_t)(x >> (32 - e)), 0xFFFFFFFFu, lo_part, c);
      uint64_t bw;
      const uint64_t w = dev::sub32(z, x1 >> (32 - e), bw);
      return dev::fold_masks(w, c, bw);
    }
  }
#else
  if constexpr (E < 32) {
    const uint64_t y = x << E;
    const uint64_t z = (uint64_t)(x1 >> (32 - E)) * EPS + y;
    return fold(z, z < y, false);
  } else {
    constexpr unsigned e = E - 32;
    const uint64_t lo_part = (uint64_t)(x0 << e) << 32;
    const uint32_t h0 = e ? (uint32_t)(x >> ((32 - e) % 32)) : x1;
    const uint32_t h1 = e ? x1 >> ((32 - e) % 32) : 0u;
    const uint64_t z = (uint64_t)h0 * EPS + lo_part;
    const bool c = z < lo_part;
    const bool bw = z < h1;
    return fold(z - h1, c && !bw, bw && !c);
  }
#endif
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

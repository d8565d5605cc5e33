// Goldilocks field arithmetic, p = 2^64 - 2^32 + 1, for host and gfx950.
//
// Semantics follow the reference's base field (native/goldilocks.cuh:90-185):
// elements are u64 words that may hold any representative in [0, 2^64);
// add/sub/mul return a representative congruent mod p, and canon() maps to
// [0, p).  The implementation is derived from the two identities
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

GL_FN uint64_t canon(uint64_t x) { return x >= P ? x - P : x; }

// a + b mod p for any u64 inputs.
GL_FN uint64_t add(uint64_t a, uint64_t b) {
  uint64_t s;
  bool c1 = __builtin_add_overflow(a, b, &s);
  // 2^64 wraps to EPS; a second wrap is possible only for non-canonical
  // inputs and then leaves s < EPS, so the third add cannot wrap.
  bool c2 = __builtin_add_overflow(s, c1 ? EPS : 0ull, &s);
  return s + (c2 ? EPS : 0ull);
}

// a - b mod p for any u64 inputs.
GL_FN uint64_t sub(uint64_t a, uint64_t b) {
  uint64_t d;
  bool b1 = __builtin_sub_overflow(a, b, &d);
  bool b2 = __builtin_sub_overflow(d, b1 ? EPS : 0ull, &d);
  return d - (b2 ? EPS : 0ull);
}

GL_FN uint64_t neg(uint64_t a) { return sub(0ull, a); }

// Reduce lo + hi * 2^64.  hi = h1 * 2^32 + h0 contributes h0 * (2^32 - 1) - h1.
GL_FN uint64_t reduce128(uint64_t lo, uint64_t hi) {
  const uint64_t h0 = hi & 0xFFFFFFFFull;
  const uint64_t h1 = hi >> 32;
  uint64_t t;
  // lo - h1: on borrow the wrapped value is >= 2^64 - 2^32, so one EPS
  // correction cannot borrow again.
  bool b = __builtin_sub_overflow(lo, h1, &t);
  t -= b ? EPS : 0ull;
  const uint64_t u = (h0 << 32) - h0;  // h0 * (2^32 - 1) < 2^64
  uint64_t r;
  bool c = __builtin_add_overflow(t, u, &r);
  // t + u - 2^64 + EPS < u + EPS < 2^64: no second wrap.
  return r + (c ? EPS : 0ull);
}

// Full 64x64 -> 128 product built from four 32x32+64 multiply-adds.
GL_FN void mul_wide(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
  const uint64_t a0 = a & 0xFFFFFFFFull, a1 = a >> 32;
  const uint64_t b0 = b & 0xFFFFFFFFull, b1 = b >> 32;
  const uint64_t p00 = a0 * b0;
  const uint64_t t1 = a0 * b1 + (p00 >> 32);          // < 2^64
  const uint64_t t2 = a1 * b0 + (t1 & 0xFFFFFFFFull);  // < 2^64
  lo = (p00 & 0xFFFFFFFFull) | (t2 << 32);
  hi = a1 * b1 + (t1 >> 32) + (t2 >> 32);               // <= 2^64 - 1
}

GL_FN uint64_t mul(uint64_t a, uint64_t b) {
  uint64_t lo, hi;
  mul_wide(a, b, lo, hi);
  return reduce128(lo, hi);
}

// Reduce lo + hi * 2^64 for hi < 2^32: hi * 2^64 = hi * (2^32 - 1).
GL_FN uint64_t reduce96(uint64_t lo, uint32_t hi) {
  const uint64_t u = ((uint64_t)hi << 32) - hi;
  uint64_t r;
  bool c = __builtin_add_overflow(lo, u, &r);
  // lo + u - 2^64 + EPS < u + EPS < 2^64: no second wrap.
  return r + (c ? EPS : 0ull);
}

// x * 2^E mod p for a compile-time exponent 0 <= E < 96.  (2 has order 192
// and 2^96 = -1; callers fold the sign of larger exponents into the
// surrounding add/sub.)
template <unsigned E> GL_FN uint64_t mul_pow2(uint64_t x) {
  static_assert(E < 96, "use the sign-folded form for exponents >= 96");
  if constexpr (E == 0) {
    return x;
  } else if constexpr (E <= 32) {
    return reduce96(x << E, (uint32_t)(x >> (64 - E)));
  } else if constexpr (E < 64) {
    return reduce128(x << E, x >> (64 - E));
  } else if constexpr (E == 64) {
    return reduce128(0ull, x);
  } else {
    // E in (64, 96): x * 2^(E-64) = lo + hi * 2^64 with hi < 2^32, then
    // times 2^64: lo * 2^64 + hi * 2^128 = lo * 2^64 - hi * 2^32.
    const uint64_t lo = x << (E - 64);
    const uint64_t hi = x >> (128 - E);
    return sub(reduce128(0ull, lo), hi << 32);
  }
}

// x * 2^E for any 0 <= E < 192.
template <unsigned E> GL_FN uint64_t mul_pow2_any(uint64_t x) {
  if constexpr (E >= 96) return neg(mul_pow2<E - 96>(x));
  else return mul_pow2<E>(x);
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

GL_FN uint64_t inv(uint64_t a) { return pow(a, P - 2); }

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

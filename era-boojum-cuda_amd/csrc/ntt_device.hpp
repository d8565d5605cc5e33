// Device helpers shared by the NTT kernel families: compile-time loops,
// power-table lookups and the in-register radix-2^A sub-DFTs.
#pragma once
#include "gl_field.hpp"
#include "ntt_internal.hpp"

#include <utility>

namespace bc {

template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F> __device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

__device__ __forceinline__ uint32_t brev_bits(uint32_t x, uint32_t bits) {
  return bits ? (__brev(x) >> (32u - bits)) : 0u;
}

// omega_{2^32}^E.
__device__ __forceinline__ uint64_t pow_w(const Tables &t, uint32_t e) {
  const uint64_t c = t.w_coarse[e >> TABLE_LOG];
  const uint32_t f = e & (TABLE_LEN - 1);
  return f ? gl::mul(c, t.w_fine[f]) : c;
}

// omega_{2^32}^E for E a multiple of 2^16.
__device__ __forceinline__ uint64_t pow_w_coarse(const Tables &t, uint32_t e) { return t.w_coarse[e >> TABLE_LOG]; }

__device__ __forceinline__ uint64_t pow_g(const Tables &t, uint32_t i, bool inverse) {
  const uint64_t *fine = inverse ? t.gi_fine : t.g_fine;
  const uint64_t *coarse = inverse ? t.gi_coarse : t.g_coarse;
  const uint64_t c = coarse[i >> TABLE_LOG];
  const uint32_t f = i & (TABLE_LEN - 1);
  return f ? gl::mul(c, fine[f]) : c;
}

// In-register DFT of 2^A points, natural input -> bit-reversed output
// (decimation in frequency).  Twiddles omega_{2^k}^t are powers of two,
// applied as sign * canonical value with the sign folded into the add/sub.
template <int A, bool INV> __device__ __forceinline__ void dft_dif(uint64_t *x) {
  static_for<A>([&](auto s) {
    constexpr int LH = A - 1 - decltype(s)::value;
    constexpr int HALF = 1 << LH;
    static_for<(1 << A) / 2>([&](auto k) {
      constexpr int K = decltype(k)::value;
      constexpr int T = K % HALF;
      constexpr int I0 = (K / HALF) * 2 * HALF + T;
      using W = gl::Pow2<gl::root_pow2_exp(LH + 1, T, INV)>;
      const uint64_t u = x[I0], v = x[I0 + HALF];
      x[I0] = gl::add(u, v);
      x[I0 + HALF] = W::value(W::NEG ? gl::sub(v, u) : gl::sub(u, v));
    });
  });
}

// In-register DFT of 2^A points, bit-reversed input -> natural output
// (decimation in time).
template <int A, bool INV> __device__ __forceinline__ void dft_dit(uint64_t *x) {
  static_for<A>([&](auto s) {
    constexpr int LH = decltype(s)::value;
    constexpr int HALF = 1 << LH;
    static_for<(1 << A) / 2>([&](auto k) {
      constexpr int K = decltype(k)::value;
      constexpr int T = K % HALF;
      constexpr int I0 = (K / HALF) * 2 * HALF + T;
      using W = gl::Pow2<gl::root_pow2_exp(LH + 1, T, INV)>;
      const uint64_t u = x[I0];
      const uint64_t m = W::value(x[I0 + HALF]);
      if constexpr (W::NEG) {
        x[I0] = gl::sub(u, m);
        x[I0 + HALF] = gl::add(u, m);
      } else {
        x[I0] = gl::add(u, m);
        x[I0 + HALF] = gl::sub(u, m);
      }
    });
  });
}

// s^i (or s^-i) for the coset shift s = g * omega_{n 2^ext}^coset, where
// coset_exp = coset << (32 - log_n - log_ext).
__device__ __forceinline__ uint64_t coset_pow(const Tables &t, uint32_t coset_exp, uint32_t i, bool inverse) {
  uint64_t v = pow_g(t, i, inverse);
  if (coset_exp) {
    uint32_t e = i * coset_exp;
    if (inverse) e = 0u - e;
    v = gl::mul(v, pow_w(t, e));
  }
  return v;
}

}  // namespace bc

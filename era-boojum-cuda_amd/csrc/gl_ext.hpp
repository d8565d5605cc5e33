// Quadratic extension of Goldilocks, F_p[u] / (u^2 - 7): the extension the
// reference's barycentric evaluation works in (native/goldilocks_extension.cuh:
// non-residue 7).  Elements are c0 + c1 u with canonical components.
#pragma once
#include "gl_field.hpp"

namespace gl {

struct E2 {
  uint64_t c0, c1;
};

constexpr uint64_t E2_NON_RESIDUE = 7;

GL_FN E2 e2_add(E2 a, E2 b) { return {add(a.c0, b.c0), add(a.c1, b.c1)}; }
GL_FN E2 e2_sub(E2 a, E2 b) { return {sub(a.c0, b.c0), sub(a.c1, b.c1)}; }

// 7 x = 8 x - x
GL_FN uint64_t mul_by_non_residue(uint64_t x) { return sub(mul_pow2<3>(x), x); }

// (a0 + a1 u)(b0 + b1 u) = a0 b0 + 7 a1 b1 + (a0 b1 + a1 b0) u
GL_FN E2 e2_mul(E2 a, E2 b) {
  const uint64_t p00 = mul(a.c0, b.c0), p11 = mul(a.c1, b.c1);
  const uint64_t p01 = mul(a.c0, b.c1), p10 = mul(a.c1, b.c0);
  return {add(p00, mul_by_non_residue(p11)), add(p01, p10)};
}

GL_FN E2 e2_mul_base(E2 a, uint64_t b) { return {mul(a.c0, b), mul(a.c1, b)}; }

// 1 / (a0 + a1 u) = (a0 - a1 u) / (a0^2 - 7 a1^2); the norm is non-zero for a
// non-zero element because 7 is not a square mod p.
GL_FN E2 e2_inv(E2 a) {
  const uint64_t norm = sub(mul(a.c0, a.c0), mul_by_non_residue(mul(a.c1, a.c1)));
  const uint64_t ni = inv(norm);
  return {mul(a.c0, ni), neg(mul(a.c1, ni))};
}

}  // namespace gl

// CPU check of the device field arithmetic (era-boojum-cuda_amd/csrc/gl_field.hpp,
// compiled here as host code) against 128-bit integer arithmetic mod
// p = 2^64 - 2^32 + 1: canon, add, sub, neg, mul, mul by 2^e for all 192
// exponents (2 has order 192), pow and inv, the unreduced multiply-accumulate
// (mac / acc_reduce) over sums of up to 4096 products, and the quadratic
// extension F_p[u]/(u^2 - 7) of gl_ext.hpp, on edge values and random draws
// biased toward carries.  Prints "OK" and exits 0 when every result matches.
#include "../../era-boojum-cuda_amd/csrc/gl_ext.hpp"

#include <stdio.h>

#include <random>
#include <utility>

typedef unsigned __int128 u128;
static const uint64_t P = gl::P;

static uint64_t ref_mul(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) % P); }
static uint64_t ref_pow(uint64_t a, uint64_t e) {
  uint64_t r = 1, b = a % P;
  for (; e; e >>= 1, b = ref_mul(b, b))
    if (e & 1) r = ref_mul(r, b);
  return r;
}

static const uint64_t EDGE[] = {0, 1, 2, P - 2, P - 1, P, P + 1, ~0ull, 0xFFFFFFFFull, 0x100000000ull, 1ull << 63,
                                P - 0x100000000ull, 0xFFFFFFFF00000000ull, 0xFFFFFFFEull, 0xFFFFFFFEFFFFFFFFull};
static const int NEDGE = sizeof(EDGE) / sizeof(EDGE[0]);

static uint64_t draw(std::mt19937_64 &g, int i) {
  if (i < NEDGE) return EDGE[i];
  switch (i % 5) {
  case 0: return g();
  case 1: return g() & 0xFFFFFFFFull;
  case 2: return g() << 32;
  case 3: return ~(g() & 0xFFFFFFFFFFull);
  default: return P - 1 - (g() & 0xFFFF);
  }
}

static int bad = 0;
static void report(const char *what, uint64_t a, uint64_t b, uint64_t got, uint64_t want) {
  if (bad++ < 8)
    printf("%s a=%016llx b=%016llx got %016llx want %016llx\n", what, (unsigned long long)a, (unsigned long long)b,
           (unsigned long long)got, (unsigned long long)want);
}

template <unsigned E> static void check_pow2(std::mt19937_64 &g, int samples) {
  const uint64_t w = ref_pow(2, E);
  for (int i = 0; i < samples; ++i) {
    uint64_t x = draw(g, i);
    if (E % 96 == 0) x %= P;  // 2^0 passes x through and 2^96 = -1 negates it: both expect canonical x
    const uint64_t got = gl::mul_pow2<E>(x), want = ref_mul(x, w);
    if (got != want) report("mul_pow2", x, E, got, want);
  }
}

template <unsigned... Es> static void check_all_pow2(std::mt19937_64 &g, int samples, std::integer_sequence<unsigned, Es...>) {
  (check_pow2<Es>(g, samples), ...);
}

int main() {
  std::mt19937_64 g(12345);
  check_all_pow2(g, 100000, std::make_integer_sequence<unsigned, 192>{});
  for (int i = 0; i < 2000000; ++i) {
    const uint64_t a = draw(g, i), b = draw(g, i / 3 + 7);
    const uint64_t ca = a % P, cb = b % P;
    if (gl::canon(a) != ca) report("canon", a, 0, gl::canon(a), ca);
    if (gl::add(ca, cb) != (uint64_t)(((u128)ca + cb) % P)) report("add", ca, cb, gl::add(ca, cb), 0);
    if (gl::sub(ca, cb) != (uint64_t)(((u128)ca + P - cb) % P)) report("sub", ca, cb, gl::sub(ca, cb), 0);
    if (gl::neg(ca) != (P - ca) % P) report("neg", ca, 0, gl::neg(ca), (P - ca) % P);
    if (gl::mul(a, b) != ref_mul(a, b)) report("mul", a, b, gl::mul(a, b), ref_mul(a, b));
  }
  if (gl::inv(0) != 0 || gl::inv(P) != 0 || gl::inv(P + 1) != 1) report("inv edge", 0, 0, gl::inv(0), 0);
  for (int i = 0; i < 2000; ++i) {
    const uint64_t a = draw(g, i) % P, e = g();
    if (gl::pow(a, e) != ref_pow(a, e)) report("pow", a, e, gl::pow(a, e), ref_pow(a, e));
    if (a && ref_mul(gl::inv(a), a) != 1) report("inv", a, 0, gl::inv(a), 0);
    if (gl::inv(a) != ref_pow(a, P - 2)) report("inv chain", a, 0, gl::inv(a), ref_pow(a, P - 2));
  }
  // mac: sums of n products of raw u64 values, n up to 4096 (carry counts
  // up to 4096 in the top limb), against the sum of reduced products.
  for (int trial = 0; trial < 3000; ++trial) {
    const int n = trial < 2000 ? 1 + trial % 40 : 4096;
    gl::Acc acc;
    uint64_t want = 0;
    for (int i = 0; i < n; ++i) {
      const uint64_t a = trial % 3 ? draw(g, trial + i) : ~0ull, b = trial % 5 ? draw(g, i * 7 + trial) : ~0ull;
      gl::mac(acc, a, b);
      want = (uint64_t)(((u128)want + ref_mul(a, b)) % P);
    }
    if (gl::acc_reduce(acc) != want) report("mac", (uint64_t)n, 0, gl::acc_reduce(acc), want);
  }
  // quadratic extension: product by schoolbook arithmetic, inverse by a * a^-1 = 1
  for (int i = 0; i < 200000; ++i) {
    const gl::E2 a{draw(g, i) % P, draw(g, i + 3) % P}, b{draw(g, i + 11) % P, draw(g, i / 2) % P};
    const gl::E2 m = gl::e2_mul(a, b);
    const uint64_t w0 = (uint64_t)(((u128)ref_mul(a.c0, b.c0) + ref_mul(7, ref_mul(a.c1, b.c1))) % P);
    const uint64_t w1 = (uint64_t)(((u128)ref_mul(a.c0, b.c1) + ref_mul(a.c1, b.c0)) % P);
    if (m.c0 != w0 || m.c1 != w1) report("e2_mul", a.c0, b.c0, m.c0, w0);
    if (gl::mul_by_non_residue(a.c1) != ref_mul(7, a.c1)) report("nr", a.c1, 0, gl::mul_by_non_residue(a.c1), 0);
    if (i < 2000 && (a.c0 | a.c1)) {
      const gl::E2 r = gl::e2_mul(gl::e2_inv(a), a);
      if (r.c0 != 1 || r.c1 != 0) report("e2_inv", a.c0, a.c1, r.c0, 1);
    }
  }
  if (bad) {
    printf("FAIL %d mismatches\n", bad);
    return 1;
  }
  printf("OK\n");
  return 0;
}

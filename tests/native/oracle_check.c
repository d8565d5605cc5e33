/* The oracle (oracle/gl_oracle.c, oracle/poseidon_oracle.c) run under
 * AddressSanitizer and UndefinedBehaviorSanitizer through every exported
 * routine the tests use, at small sizes, with checks that fix the results:
 * round trips, the O(n^2) DFT, Horner evaluation of cosets, the threaded
 * batch against the serial one, barycentric evaluation against Horner, and the
 * fused Poseidon permutation against the textbook one.  Prints "OK" on
 * success. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

uint64_t gl_add(uint64_t a, uint64_t b);
uint64_t gl_mul(uint64_t a, uint64_t b);
uint64_t gl_pow(uint64_t b, uint64_t e);
uint64_t gl_domain_generator(uint32_t log_size);
uint64_t gl_bitrev(uint64_t x, uint32_t bits);
uint64_t gl_coset_shift(uint32_t log_n, uint32_t log_ext, uint32_t coset_index);
void gl_ntt_column(const uint64_t *in, uint64_t *out, uint32_t log_n, int bitrev_inputs, int inverse,
                   uint32_t log_ext, uint32_t coset_index, const uint64_t *tw);
void gl_batch_ntt(const uint64_t *in, uint64_t *out, uint32_t log_n, uint32_t num_ntts, uint64_t stride_in,
                  uint64_t stride_out, int bitrev_inputs, int inverse, uint32_t log_ext, uint32_t coset_index);
int gl_batch_ntt_mt(const uint64_t *in, uint64_t *out, uint32_t log_n, uint32_t num_ntts, uint64_t stride_in,
                    uint64_t stride_out, int bitrev_inputs, int inverse, uint32_t log_ext, uint32_t coset_index,
                    uint32_t threads);
void gl_dft_naive(const uint64_t *a, uint64_t *out, uint64_t n, uint64_t root);
uint64_t gl_horner(const uint64_t *a, uint64_t n, uint64_t x);
void gl_fill_uniform(uint64_t *out, uint64_t n, uint64_t seed, uint64_t column);
void gl_fill_raw(uint64_t *out, uint64_t n, uint64_t seed, uint64_t column);
void gl_bary_lagrange_coeffs(const uint64_t *x, int ext, uint64_t coset, uint32_t log_count, int bit_reverse,
                             uint64_t *coeffs_c0, uint64_t *coeffs_c1);
void gl_bary_batch_eval(const uint64_t *ys, int ys_ext, const uint64_t *coeffs, int coef_ext, uint64_t count,
                        uint32_t num_polys, uint64_t *evals);
void gl_e2_horner(const uint64_t *a0, const uint64_t *a1, uint64_t n, const uint64_t *x, uint64_t *out);
size_t pos_fused_words(void);
int pos_derive(const uint64_t *rc, uint64_t *out);
void pos_permute_textbook(uint64_t *s, const uint64_t *rc);
void pos_permute_fused(uint64_t *s, const uint64_t *rc, const uint64_t *fused);
void pos_leaves(const uint64_t *values, uint64_t *results, uint32_t rows, uint32_t cols, uint32_t count,
                int load_intermediate, int store_intermediate, const uint64_t *rc, const uint64_t *fused);
void pos_nodes(const uint64_t *values, uint64_t *results, uint32_t count, const uint64_t *rc, const uint64_t *fused);
void pos_grain_constants(uint64_t *out);
void pos2_permute(uint64_t *s, const uint64_t *rc);
void pos2_leaves(const uint64_t *values, uint64_t *results, uint32_t rows, uint32_t cols, uint32_t count,
                 int load_intermediate, int store_intermediate, const uint64_t *rc);
void pos2_nodes(const uint64_t *values, uint64_t *results, uint32_t count, const uint64_t *rc);

#define REQUIRE(c)                                                  \
  do {                                                              \
    if (!(c)) {                                                     \
      printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);           \
      exit(1);                                                      \
    }                                                               \
  } while (0)

int main(void) {
  /* transforms: round trips for every order and coset, DFT and Horner */
  for (uint32_t log_n = 1; log_n <= 11; ++log_n) {
    const uint64_t n = 1ull << log_n;
    uint64_t *x = malloc(n * 8), *y = malloc(n * 8), *z = malloc(n * 8), *d = malloc(n * 8);
    gl_fill_raw(x, n, 5, log_n);
    uint64_t *xc = malloc(n * 8);
    for (uint64_t i = 0; i < n; ++i) xc[i] = gl_add(x[i], 0);  /* canonical copy */
    for (int br = 0; br < 2; ++br)
      for (uint32_t k = 0; k <= 3 && log_n + k <= 14; ++k)
        for (uint32_t c = 0; c < (1u << k); c += (k ? (1u << k) - 1 : 1)) {
          gl_ntt_column(x, y, log_n, br, 0, k, c, NULL);
          gl_ntt_column(y, z, log_n, !br, 1, k, c, NULL);
          REQUIRE(memcmp(z, xc, n * 8) == 0);
        }
    if (log_n <= 8) {
      gl_ntt_column(x, y, log_n, 0, 0, 0, 0, NULL);  /* natural in, bit-reversed out */
      gl_dft_naive(xc, d, n, gl_domain_generator(log_n));
      for (uint64_t j = 0; j < n; ++j) REQUIRE(y[gl_bitrev(j, log_n)] == d[j]);
      gl_ntt_column(x, y, log_n, 0, 0, 2, 3, NULL);  /* coset 3 of blowup 4 */
      const uint64_t s = gl_coset_shift(log_n, 2, 3), w = gl_domain_generator(log_n);
      for (uint64_t j = 0; j < n; j += 7) REQUIRE(y[gl_bitrev(j, log_n)] == gl_horner(xc, n, gl_mul(s, gl_pow(w, j))));
    }
    free(x); free(y); free(z); free(d); free(xc);
  }
  /* threaded batch with padded strides equals the serial one */
  {
    const uint32_t log_n = 10, cols = 9;
    const uint64_t n = 1ull << log_n, si = n + 3, so = n + 5;
    uint64_t *in = calloc(cols * si, 8), *a = calloc(cols * so, 8), *b = calloc(cols * so, 8);
    for (uint32_t c = 0; c < cols; ++c) gl_fill_uniform(in + c * si, n, 7, c);
    gl_batch_ntt(in, a, log_n, cols, si, so, 1, 1, 0, 0);
    REQUIRE(gl_batch_ntt_mt(in, b, log_n, cols, si, so, 1, 1, 0, 0, 4) == 0);
    REQUIRE(memcmp(a, b, cols * so * 8) == 0);
    free(in); free(a); free(b);
  }
  /* barycentric evaluation at a point off the domain equals Horner on the
   * coefficients of the interpolant */
  {
    const uint32_t log_n = 6;
    const uint64_t n = 1ull << log_n, coset = 7;
    uint64_t *coef = malloc(2 * n * 8), *ys = malloc(2 * n * 8), *l0 = malloc(n * 8), *l1 = malloc(n * 8);
    gl_fill_uniform(coef, n, 9, 0);
    gl_fill_uniform(coef + n, n, 9, 1);
    const uint64_t w = gl_domain_generator(log_n);
    for (uint64_t i = 0; i < n; ++i) {
      const uint64_t pt = gl_mul(coset, gl_pow(w, i));
      ys[i] = gl_horner(coef, n, pt);
      ys[n + i] = gl_horner(coef + n, n, pt);
    }
    const uint64_t x[2] = {123456789, 987654321};
    gl_bary_lagrange_coeffs(x, 1, coset, log_n, 0, l0, l1);
    uint64_t lc[2 * 64], ev[2], want[2];
    memcpy(lc, l0, n * 8);
    memcpy(lc + n, l1, n * 8);
    gl_bary_batch_eval(ys, 1, lc, 1, n, 1, ev);
    gl_e2_horner(coef, coef + n, n, x, want);
    REQUIRE(ev[0] == want[0] && ev[1] == want[1]);
    free(coef); free(ys); free(l0); free(l1);
  }
  /* Poseidon: fused form equals the textbook permutation; leaves and nodes;
   * Poseidon2 leaves and nodes run */
  {
    uint64_t rc[360];
    pos_grain_constants(rc);
    const size_t words = pos_fused_words();
    uint64_t *fused = malloc(words * 8);
    REQUIRE(pos_derive(rc, fused) == 1);
    for (int t = 0; t < 4; ++t) {
      uint64_t s1[12], s2[12];
      gl_fill_raw(s1, 12, 11, t);
      for (int i = 0; i < 12; ++i) s1[i] = gl_add(s1[i], 0);
      memcpy(s2, s1, sizeof s1);
      pos_permute_textbook(s1, rc);
      pos_permute_fused(s2, rc, fused);
      REQUIRE(memcmp(s1, s2, sizeof s1) == 0);
    }
    const uint32_t rows = 4, cols = 5, count = 8;
    uint64_t *vals = malloc((size_t)rows * cols * count * 8), res[4 * 8], nodes[4 * 4], res2[4 * 8], nodes2[4 * 4];
    gl_fill_raw(vals, (uint64_t)rows * cols * count, 13, 0);
    pos_leaves(vals, res, rows, cols, count, 0, 0, rc, fused);
    pos_nodes(res, nodes, count / 2, rc, fused);
    pos2_leaves(vals, res2, rows, cols, count, 0, 0, rc);
    pos2_nodes(res2, nodes2, count / 2, rc);
    /* a leaf of one column chunk of 8 values is one permutation of that chunk */
    uint64_t s[12] = {0};
    for (int i = 0; i < 8; ++i) s[i] = vals[i]; /* count 1: value (r, c) at r + c rows */
    uint64_t one[4];
    pos_leaves(vals, one, rows, 2, 1, 0, 0, rc, fused);
    for (int i = 0; i < 8; ++i) s[i] = gl_add(s[i], 0);
    pos_permute_textbook(s, rc);
    for (int i = 0; i < 4; ++i) REQUIRE(one[i] == s[i]);
    free(vals); free(fused);
  }
  printf("OK\n");
  return 0;
}

/*
 * CPU oracle for the batched Goldilocks NTT / iNTT / coset-LDE path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker or the CPU baseline.
 *
 * What it restates.  The reference checks its GPU NTT against the CPU FFT of
 * the third-party crate `boojum = "=0.2.2"` (pinned at Cargo.toml:21 and
 * Cargo.lock:84-87 of the reference; its source is not vendored).  The call
 * sites that fix the semantics are:
 *   - src/ntt.rs:404-416   precompute_twiddles_for_fft::<.., INVERSED>(n_max)
 *   - src/ntt.rs:441-447   bitreverse_enumeration_inplace on the b2n inputs
 *   - src/ntt.rs:527-537   coset = g * omega_{n*2^k}^c when log_ext > 0
 *   - src/ntt.rs:575-583   fft_natural_to_bitreversed / ifft_natural_to_natural
 *   - src/context.rs:182-252  omega / g / inv_sizes tables (Context::create)
 * The boojum algorithms restated here (from their published behaviour):
 *   - domain_generator_for_size(M) = radix-2 subgroup generator squared
 *     (32 - log2 M) times;
 *   - precompute_twiddles_for_fft(n, inv) = bit-reversed first n/2 powers of
 *     omega_n (or omega_n^-1);
 *   - distribute_powers(a, b): a[i] *= b^i;
 *   - serial Cooley-Tukey natural->bit-reversed with bit-reversed twiddles;
 *   - fft_natural_to_bitreversed(a, coset, tw) = distribute_powers(coset) then
 *     the serial CT transform;
 *   - ifft_natural_to_natural(a, coset, tw) = serial CT (inverse twiddles),
 *     bit-reverse, distribute_powers(coset^-1), scale by n^-1.
 *
 * Parity pinning (tests/test_oracle.py): field ops against Python big-int
 * arithmetic; the two constants (g = 7 generates F_p^*, omega_{2^32} =
 * 7^((p-1)/2^32) has exact order 2^32); every transform against an O(n^2)
 * DFT / Horner evaluation written independently in Python; round trips;
 * linearity; delta -> all-ones; the LDE interleave identity.
 *
 * Arithmetic is textbook: 128-bit products reduced with '%'.  Outputs are
 * canonical in [0, p).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

typedef unsigned __int128 u128;

#define GL_P 0xFFFFFFFF00000001ull

static inline uint64_t gl_red(uint64_t a) { return a >= GL_P ? a - GL_P : a; }

uint64_t gl_add(uint64_t a, uint64_t b) {
  u128 s = (u128)gl_red(a) + gl_red(b);
  return (uint64_t)(s % GL_P);
}

uint64_t gl_sub(uint64_t a, uint64_t b) {
  u128 s = (u128)gl_red(a) + GL_P - gl_red(b);
  return (uint64_t)(s % GL_P);
}

uint64_t gl_mul(uint64_t a, uint64_t b) {
  u128 s = (u128)a * b;
  return (uint64_t)(s % GL_P);
}

uint64_t gl_pow(uint64_t b, uint64_t e) {
  uint64_t r = 1;
  b = gl_red(b);
  while (e) {
    if (e & 1) r = gl_mul(r, b);
    b = gl_mul(b, b);
    e >>= 1;
  }
  return r;
}

uint64_t gl_inv(uint64_t a) { return gl_pow(a, GL_P - 2); }

/* boojum GoldilocksField::multiplicative_generator(); used at
 * src/context.rs:232 and src/ntt.rs:535 of the reference. */
uint64_t gl_generator(void) { return 7; }

/* Radix-2 subgroup generator of order 2^32: 7^((p-1)/2^32). */
uint64_t gl_omega_2_32(void) { return gl_pow(7, (GL_P - 1) >> 32); }

/* boojum domain_generator_for_size(2^log_size); reference call sites
 * src/ntt.rs:530 and src/context.rs:182-219. */
uint64_t gl_domain_generator(uint32_t log_size) {
  uint64_t w = gl_omega_2_32();
  for (uint32_t i = log_size; i < 32; i++) w = gl_mul(w, w);
  return w;
}

uint64_t gl_bitrev(uint64_t x, uint32_t bits) {
  uint64_t r = 0;
  for (uint32_t i = 0; i < bits; i++) r |= ((x >> i) & 1ull) << (bits - 1 - i);
  return r;
}

static uint32_t ilog2_u64(uint64_t n) {
  uint32_t l = 0;
  while ((1ull << l) < n) l++;
  return l;
}

/* boojum bitreverse_enumeration_inplace; reference call sites
 * src/ntt.rs:446,577-582 and src/context.rs:152. */
void gl_bitreverse_inplace(uint64_t *a, uint64_t n) {
  uint32_t bits = ilog2_u64(n);
  for (uint64_t i = 0; i < n; i++) {
    uint64_t j = gl_bitrev(i, bits);
    if (j > i) {
      uint64_t t = a[i];
      a[i] = a[j];
      a[j] = t;
    }
  }
}

/* boojum distribute_powers(a, base): a[i] *= base^i; reference call sites
 * src/context.rs:150,252. */
void gl_distribute_powers(uint64_t *a, uint64_t n, uint64_t base) {
  uint64_t p = 1;
  for (uint64_t i = 0; i < n; i++) {
    a[i] = gl_mul(a[i], p);
    p = gl_mul(p, base);
  }
}

/* boojum precompute_twiddles_for_fft::<.., INVERSED>(n): the first n/2
 * powers of omega_n (inverted when `inverse`), bit-reversed.  Reference call
 * site src/ntt.rs:404-416. */
void gl_precompute_twiddles(uint64_t n, int inverse, uint64_t *out) {
  uint64_t half = n / 2;
  if (half == 0) return;
  uint64_t w = gl_domain_generator(ilog2_u64(n));
  if (inverse) w = gl_inv(w);
  uint64_t p = 1;
  for (uint64_t i = 0; i < half; i++) {
    out[i] = p;
    p = gl_mul(p, w);
  }
  gl_bitreverse_inplace(out, half);
}

/* Serial Cooley-Tukey natural -> bit-reversed transform with bit-reversed
 * twiddles (boojum's serial_ct_ntt_natural_to_bitreversed). */
void gl_serial_ct_n2b(uint64_t *a, uint64_t n, const uint64_t *tw_bitrev) {
  if (n <= 1) return;
  uint64_t pairs_per_group = n / 2, num_groups = 1, distance = n / 2;
  while (num_groups < n) {
    for (uint64_t k = 0; k < num_groups; k++) {
      uint64_t i1 = k * pairs_per_group * 2;
      uint64_t i2 = i1 + pairs_per_group;
      uint64_t s = tw_bitrev[k];
      for (uint64_t j = i1; j < i2; j++) {
        uint64_t u = a[j];
        uint64_t v = gl_mul(a[j + distance], s);
        a[j + distance] = gl_sub(u, v);
        a[j] = gl_add(u, v);
      }
    }
    pairs_per_group /= 2;
    num_groups *= 2;
    distance /= 2;
  }
}

/* boojum fft_natural_to_bitreversed(input, coset, twiddles); reference call
 * site src/ntt.rs:580. */
void gl_fft_natural_to_bitreversed(uint64_t *a, uint64_t n, uint64_t coset, const uint64_t *tw) {
  for (uint64_t i = 0; i < n; i++) a[i] = gl_red(a[i]);
  if (coset != 1) gl_distribute_powers(a, n, coset);
  gl_serial_ct_n2b(a, n, tw);
}

/* boojum ifft_natural_to_natural(input, coset, twiddles); reference call
 * site src/ntt.rs:576. */
void gl_ifft_natural_to_natural(uint64_t *a, uint64_t n, uint64_t coset, const uint64_t *tw_inv) {
  for (uint64_t i = 0; i < n; i++) a[i] = gl_red(a[i]);
  gl_serial_ct_n2b(a, n, tw_inv);
  gl_bitreverse_inplace(a, n);
  if (coset != 1) gl_distribute_powers(a, n, gl_inv(coset));
  if (n > 1) {
    uint64_t n_inv = gl_inv(n % GL_P);
    for (uint64_t i = 0; i < n; i++) a[i] = gl_mul(a[i], n_inv);
  }
}

/* Coset shift used by the reference test (src/ntt.rs:527-537):
 * 1 when log_ext == 0, else g * omega_{n*2^log_ext}^coset_index. */
uint64_t gl_coset_shift(uint32_t log_n, uint32_t log_ext, uint32_t coset_index) {
  if (log_ext == 0) return 1;
  uint64_t w = gl_domain_generator(log_n + log_ext);
  return gl_mul(gl_generator(), gl_pow(w, coset_index));
}

/*
 * One column of the reference's batch NTT, in the layout the GPU kernels
 * produce (src/ntt.rs:441-447 and 575-583 define the order conventions):
 *   bitrev_inputs = 0: natural input, output bit-reversed;
 *   bitrev_inputs = 1: bit-reversed input, output natural.
 * `in` and `out` may alias.  `tw` must hold precompute_twiddles(n, inverse)
 * (or a longer prefix-compatible table, as at src/ntt.rs:401-403); pass NULL
 * to have it computed here.
 */
void gl_ntt_column(const uint64_t *in, uint64_t *out, uint32_t log_n, int bitrev_inputs, int inverse,
                   uint32_t log_ext, uint32_t coset_index, const uint64_t *tw) {
  uint64_t n = 1ull << log_n;
  uint64_t *own = NULL;
  if (!tw) {
    own = (uint64_t *)malloc(sizeof(uint64_t) * (n / 2 + 1));
    gl_precompute_twiddles(n, inverse, own);
    tw = own;
  }
  uint64_t *v = (uint64_t *)malloc(sizeof(uint64_t) * n);
  memcpy(v, in, sizeof(uint64_t) * n);
  if (bitrev_inputs) gl_bitreverse_inplace(v, n);
  uint64_t coset = gl_coset_shift(log_n, log_ext, coset_index);
  if (inverse) {
    gl_ifft_natural_to_natural(v, n, coset, tw);
  } else {
    gl_fft_natural_to_bitreversed(v, n, coset, tw);
  }
  /* natural-order result of the inverse is bit-reversed for n2b; the
   * bit-reversed result of the forward is put in natural order for b2n. */
  if (inverse != bitrev_inputs) {
    /* inverse && !bitrev: n2b inverse -> bit-reversed output.
     * !inverse && bitrev: b2n forward -> natural output. */
    gl_bitreverse_inplace(v, n);
  }
  memcpy(out, v, sizeof(uint64_t) * n);
  free(v);
  free(own);
}

/* Batched form over a column-major matrix with explicit strides. */
void gl_batch_ntt(const uint64_t *in, uint64_t *out, uint32_t log_n, uint32_t num_ntts, uint64_t stride_in,
                  uint64_t stride_out, int bitrev_inputs, int inverse, uint32_t log_ext, uint32_t coset_index) {
  uint64_t n = 1ull << log_n;
  uint64_t *tw = (uint64_t *)malloc(sizeof(uint64_t) * (n / 2 + 1));
  gl_precompute_twiddles(n, inverse, tw);
  for (uint32_t c = 0; c < num_ntts; c++)
    gl_ntt_column(in + (uint64_t)c * stride_in, out + (uint64_t)c * stride_out, log_n, bitrev_inputs, inverse, log_ext,
                  coset_index, tw);
  free(tw);
}

/* ---------------------------------------------------------------------- */
/* Multithreaded CPU baseline: the same restatement, one column per task.  */

typedef struct {
  const uint64_t *in;
  uint64_t *out;
  uint32_t log_n, num_ntts;
  uint64_t stride_in, stride_out;
  int bitrev_inputs, inverse;
  uint32_t log_ext, coset_index;
  const uint64_t *tw;
  volatile uint32_t next;
  pthread_mutex_t mu;
} mt_job;

static void *mt_worker(void *arg) {
  mt_job *j = (mt_job *)arg;
  for (;;) {
    pthread_mutex_lock(&j->mu);
    uint32_t c = j->next++;
    pthread_mutex_unlock(&j->mu);
    if (c >= j->num_ntts) break;
    gl_ntt_column(j->in + (uint64_t)c * j->stride_in, j->out + (uint64_t)c * j->stride_out, j->log_n,
                  j->bitrev_inputs, j->inverse, j->log_ext, j->coset_index, j->tw);
  }
  return NULL;
}

int gl_batch_ntt_mt(const uint64_t *in, uint64_t *out, uint32_t log_n, uint32_t num_ntts, uint64_t stride_in,
                    uint64_t stride_out, int bitrev_inputs, int inverse, uint32_t log_ext, uint32_t coset_index,
                    uint32_t threads) {
  uint64_t n = 1ull << log_n;
  uint64_t *tw = (uint64_t *)malloc(sizeof(uint64_t) * (n / 2 + 1));
  gl_precompute_twiddles(n, inverse, tw);
  mt_job j;
  j.in = in; j.out = out; j.log_n = log_n; j.num_ntts = num_ntts;
  j.stride_in = stride_in; j.stride_out = stride_out;
  j.bitrev_inputs = bitrev_inputs; j.inverse = inverse;
  j.log_ext = log_ext; j.coset_index = coset_index; j.tw = tw; j.next = 0;
  pthread_mutex_init(&j.mu, NULL);
  if (threads < 1) threads = 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
  uint32_t started = 0;
  for (uint32_t t = 0; t < threads; t++)
    if (pthread_create(&th[t], NULL, mt_worker, &j) == 0) started++;
  if (started == 0) mt_worker(&j);
  for (uint32_t t = 0; t < started; t++) pthread_join(th[t], NULL);
  pthread_mutex_destroy(&j.mu);
  free(th);
  free(tw);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Independent checks used to pin the restatement.                         */

/* O(n^2) DFT: out[k] = sum_i a[i] * root^(i*k). */
void gl_dft_naive(const uint64_t *a, uint64_t *out, uint64_t n, uint64_t root) {
  for (uint64_t k = 0; k < n; k++) {
    uint64_t step = gl_pow(root, k), p = 1, acc = 0;
    for (uint64_t i = 0; i < n; i++) {
      acc = gl_add(acc, gl_mul(a[i], p));
      p = gl_mul(p, step);
    }
    out[k] = acc;
  }
}

/* Horner evaluation of sum_i a[i] x^i. */
uint64_t gl_horner(const uint64_t *a, uint64_t n, uint64_t x) {
  uint64_t acc = 0;
  for (uint64_t i = n; i-- > 0;) acc = gl_add(gl_mul(acc, x), a[i]);
  return acc;
}

/* ---------------------------------------------------------------------- */
/* Seeded input generator: splitmix64 per column, values >= p rejected, so
 * the draws are uniform in [0, p).  Column c of a run with seed s uses the
 * stream seeded with s * 0x9E3779B97F4A7C15 + c. */

static inline uint64_t splitmix64(uint64_t *state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void gl_fill_uniform(uint64_t *out, uint64_t n, uint64_t seed, uint64_t column) {
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + column;
  for (uint64_t i = 0; i < n; i++) {
    uint64_t v;
    do v = splitmix64(&st);
    while (v >= GL_P);
    out[i] = v;
  }
}

/* Raw u64 draws (may be non-canonical), as the reference's tests use
 * (src/ntt.rs:424 draws any u64). */
void gl_fill_raw(uint64_t *out, uint64_t n, uint64_t seed, uint64_t column) {
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + column;
  for (uint64_t i = 0; i < n; i++) out[i] = splitmix64(&st);
}

/* ---------------------------------------------------------------------- */
/* Quadratic extension F_p[u]/(u^2 - 7) and barycentric evaluation.        */
/*
 * Restates the reference's native/barycentric.cu and src/barycentric.rs:
 *   - common factor (barycentric.cu:12-21):
 *       common = coset * (x^N - coset^N) / (N * coset^N)
 *   - Lagrange coefficient i (barycentric.cu:44-86):
 *       L_i = common / (x * omega_N^-i - coset), natural order i; with
 *       bit_reverse the vector is then bit-reversed (src/barycentric.rs:117-124)
 *   - batch_eval (barycentric.cu:104-156 partial sums, then the per-column
 *     sum of src/barycentric.rs:249-260): eval_p = sum_i L_i * y_{i,p}.
 * Extension elements are (c0, c1) = c0 + c1 u with u^2 = 7
 * (native/goldilocks_extension.cuh, non-residue 7).  Layouts are the
 * reference's: x and the common factor are one element (c0, c1 adjacent);
 * coefficient vectors and extension matrices are "double vectors" (the c0
 * plane of a column, then its c1 plane); extension evals are (c0, c1) pairs.
 * Pinned by tests/test_oracle.py: Ext2 ops against Python big-int
 * polynomial arithmetic, sum_i L_i(x) = 1, and batch_eval of coset
 * evaluations against Horner's rule on the monomial coefficients.
 */

void gl_e2_mul(const uint64_t *a, const uint64_t *b, uint64_t *out) {
  const uint64_t p00 = gl_mul(a[0], b[0]), p11 = gl_mul(a[1], b[1]);
  const uint64_t c0 = gl_add(p00, gl_mul(7, p11));
  const uint64_t c1 = gl_add(gl_mul(a[0], b[1]), gl_mul(a[1], b[0]));
  out[0] = c0;
  out[1] = c1;
}

void gl_e2_inv(const uint64_t *a, uint64_t *out) {
  const uint64_t norm = gl_sub(gl_mul(a[0], a[0]), gl_mul(7, gl_mul(a[1], a[1])));
  const uint64_t ni = gl_inv(norm);
  out[0] = gl_mul(a[0], ni);
  out[1] = gl_sub(0, gl_mul(a[1], ni));
}

/* Canonical copy of an element given as any u64 (reference fields accept
 * raw representatives). */
static uint64_t gl_canon(uint64_t a) { return gl_red(a); }

void gl_bary_common_factor(const uint64_t *x, int ext, uint64_t coset, uint64_t count, uint64_t *out) {
  coset = gl_canon(coset);
  const uint64_t coset_n = gl_pow(coset, count);
  const uint64_t denom_inv = gl_inv(gl_mul(gl_red(count), coset_n));
  const uint64_t f = gl_mul(coset, denom_inv);
  if (!ext) {
    out[0] = gl_mul(gl_sub(gl_pow(gl_canon(x[0]), count), coset_n), f);
    return;
  }
  uint64_t xn[2] = {1, 0}, b[2] = {gl_canon(x[0]), gl_canon(x[1])};
  for (uint64_t e = count; e; e >>= 1) { /* square and multiply */
    if (e & 1) gl_e2_mul(xn, b, xn);
    gl_e2_mul(b, b, b);
  }
  out[0] = gl_mul(gl_sub(xn[0], coset_n), f);
  out[1] = gl_mul(xn[1], f);
}

/* coeffs_c1 is ignored (may be NULL) when !ext. */
void gl_bary_lagrange_coeffs(const uint64_t *x, int ext, uint64_t coset, uint32_t log_count, int bit_reverse,
                             uint64_t *coeffs_c0, uint64_t *coeffs_c1) {
  const uint64_t count = 1ull << log_count;
  uint64_t common[2] = {0, 0};
  gl_bary_common_factor(x, ext, coset, count, common);
  coset = gl_canon(coset);
  const uint64_t w_inv = gl_inv(gl_domain_generator(log_count));
  const uint64_t x0 = gl_canon(x[0]), x1 = ext ? gl_canon(x[1]) : 0;
  /* denominators d_i = x * w^-i - coset, inverted with one field inversion
   * (prefix products, Montgomery's trick) */
  uint64_t *d0 = malloc(count * sizeof(uint64_t)), *d1 = malloc(count * sizeof(uint64_t));
  uint64_t *pre0 = malloc(count * sizeof(uint64_t)), *pre1 = malloc(count * sizeof(uint64_t));
  uint64_t w = 1, acc[2] = {1, 0};
  for (uint64_t i = 0; i < count; i++) {
    d0[i] = gl_sub(gl_mul(x0, w), coset);
    d1[i] = gl_mul(x1, w);
    pre0[i] = acc[0];
    pre1[i] = acc[1];
    const uint64_t di[2] = {d0[i], d1[i]};
    gl_e2_mul(acc, di, acc);
    w = gl_mul(w, w_inv);
  }
  uint64_t inv_acc[2];
  gl_e2_inv(acc, inv_acc);
  for (uint64_t i = count; i-- > 0;) {
    const uint64_t pi[2] = {pre0[i], pre1[i]}, di[2] = {d0[i], d1[i]};
    uint64_t inv_i[2], l[2];
    gl_e2_mul(inv_acc, pi, inv_i); /* 1 / d_i */
    gl_e2_mul(inv_acc, di, inv_acc); /* 1 / (d_0 ... d_{i-1}) */
    gl_e2_mul(inv_i, common, l);
    const uint64_t dst = bit_reverse ? gl_bitrev(i, log_count) : i;
    coeffs_c0[dst] = l[0];
    if (ext) coeffs_c1[dst] = l[1];
  }
  free(d0);
  free(d1);
  free(pre0);
  free(pre1);
}

/* evals: num_polys values (base) or (c0, c1) pairs when ys_ext or coef_ext.
 * Base ys: column p at ys + p * count.  Extension ys: column p's c0 plane at
 * ys + 2 p count, its c1 plane right after.  Extension coefficients: c0 plane
 * then c1 plane (count each). */
void gl_bary_batch_eval(const uint64_t *ys, int ys_ext, const uint64_t *coeffs, int coef_ext, uint64_t count,
                        uint32_t num_polys, uint64_t *evals) {
  const int out_ext = ys_ext || coef_ext;
  for (uint32_t p = 0; p < num_polys; p++) {
    const uint64_t *y0 = ys + (uint64_t)p * count * (ys_ext ? 2 : 1);
    const uint64_t *y1 = ys_ext ? y0 + count : NULL;
    uint64_t acc[2] = {0, 0};
    for (uint64_t i = 0; i < count; i++) {
      const uint64_t l[2] = {coeffs[i], coef_ext ? coeffs[count + i] : 0};
      const uint64_t y[2] = {gl_canon(y0[i]), y1 ? gl_canon(y1[i]) : 0};
      uint64_t t[2];
      gl_e2_mul(l, y, t);
      acc[0] = gl_add(acc[0], t[0]);
      acc[1] = gl_add(acc[1], t[1]);
    }
    if (out_ext) {
      evals[2 * p] = acc[0];
      evals[2 * p + 1] = acc[1];
    } else {
      evals[p] = acc[0];
    }
  }
}

/* Horner's rule in the extension: sum_i (a0[i] + a1[i] u) x^i (a1 may be
 * NULL for base coefficients). */
void gl_e2_horner(const uint64_t *a0, const uint64_t *a1, uint64_t n, const uint64_t *x, uint64_t *out) {
  uint64_t acc[2] = {0, 0};
  for (uint64_t i = n; i-- > 0;) {
    gl_e2_mul(acc, x, acc);
    acc[0] = gl_add(acc[0], gl_canon(a0[i]));
    acc[1] = gl_add(acc[1], a1 ? gl_canon(a1[i]) : 0);
  }
  out[0] = acc[0];
  out[1] = acc[1];
}

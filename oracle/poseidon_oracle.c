/*
 * CPU oracle for Poseidon (Goldilocks, width 12, rate 8, capacity 4, 8 full
 * and 22 partial rounds, x^7 S-box) Merkle-tree hashing.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only tests/ and bench.py use it, as the checker.
 *
 * What it restates (reference files):
 *   - native/poseidon_single_thread.cu:14-32  the permutation as the kernels
 *     run it: full rounds 0-3, with round 3's MDS fused with the first
 *     partial round's dense matrix; 22 partial rounds in sparse form; full
 *     rounds 26-29, round 26 without its constants (folded into the partial
 *     rounds);
 *   - native/poseidon_single_thread.cuh:18-145  round constants, S-box x^7,
 *     MDS = circulant of 2^MDS_MATRIX_EXPS (poseidon_constants_template.cuh:
 *     26), fused constant layer, fused dense matrix, sparse partial round;
 *   - native/poseidon_utils.cuh:8-61  leaf absorption (overwrite mode, rate
 *     8, zero padding of the last chunk, capacity checkpoints) and node
 *     hashing (two 4-element digests -> one);
 *   - src/poseidon.rs:129-170, 218-238, 262-328  argument checks and the
 *     Merkle tree layer layout.
 *
 * The round constants.  The reference takes boojum's 360 textbook round
 * constants (ALL_ROUND_CONSTANTS of boojum::implementations::
 * poseidon_goldilocks_params, crate boojum = "=0.2.2") at build time
 * (build/poseidon_constants.rs:66-75) and derives the fused form from them.
 * Those constants are not in the reference tree, so this oracle takes the 360
 * constants as an argument and derives the fused form itself
 * (pos_derive): moving the partial rounds' constants backwards through
 * M^-1 and factoring each partial round's MDS as B_k A_k, B_k sparse and A_k
 * = diag(1, N^), where N starts at M and becomes A_k M for the round before.
 *
 * Poseidon2 (pos2_*, at the end of this file) restates the reference's
 * second hash with the same constants, in textbook matrix form.
 *
 * Parity pinning (tests/test_oracle_poseidon.py):
 *   - the derived fused matrices (dense, VS, W_HATS), which depend only on the
 *     MDS, equal the reference's tables in poseidon_constants_template.cuh
 *     (checksums and spot values in tests/golden/poseidon_fused.json);
 *   - the fused permutation equals the textbook one (constants before each
 *     round's S-box, full / partial S-box layer, MDS) for any constants;
 *   - leaf and node hashing against an independent Python sponge;
 *   - Poseidon2: M4 entries equal the paper's, mu equals the reference's
 *     LOG_MU_MINUS_ONE (fixture), both matrices invertible mod p, the
 *     permutation against a Python restatement built tile by tile.
 * Parity unpinned: the constant values themselves (boojum's), hence every
 * digest computed with real boojum constants.  pos_grain_constants
 * generates constants by the Poseidon reference parameter procedure (Grain
 * LFSR, self-shrinking, rejection of values >= p); they are only test
 * inputs, and nothing shows that boojum's are these.
 *
 * Arithmetic is textbook: 128-bit products reduced with '%'.
 */
#include <stdint.h>
#include <string.h>

#define PO_P 0xFFFFFFFF00000001ull
#define PO_T 12
#define PO_RATE 8
#define PO_CAP 4
#define PO_ROUNDS 30
#define PO_HALF_FULL 4
#define PO_PARTIAL 22

static const unsigned MDS_EXPS[PO_T] = {0, 0, 1, 0, 3, 5, 1, 8, 12, 3, 16, 10};

static uint64_t po_add(uint64_t a, uint64_t b) {
  unsigned __int128 s = (unsigned __int128)(a % PO_P) + (b % PO_P);
  return (uint64_t)(s % PO_P);
}
static uint64_t po_sub(uint64_t a, uint64_t b) { return po_add(a % PO_P, PO_P - b % PO_P); }
static uint64_t po_mul(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) % PO_P); }
static uint64_t po_pow(uint64_t b, uint64_t e) {
  uint64_t r = 1;
  b %= PO_P;
  while (e) {
    if (e & 1) r = po_mul(r, b);
    b = po_mul(b, b);
    e >>= 1;
  }
  return r;
}
static uint64_t po_sbox(uint64_t x) { return po_pow(x, 7); }

/* MDS entry M[r][c] = 2^MDS_EXPS[(c - r) mod 12], so y_r = sum_i 2^e_i x_(r+i)
 * (native/poseidon_single_thread.cuh:84-104). */
static uint64_t mds_entry(int r, int c) { return 1ull << MDS_EXPS[((c - r) % PO_T + PO_T) % PO_T]; }

static void mat_vec(const uint64_t *m, int n, const uint64_t *v, uint64_t *out) {
  uint64_t t[PO_T];
  for (int r = 0; r < n; ++r) {
    uint64_t acc = 0;
    for (int c = 0; c < n; ++c) acc = po_add(acc, po_mul(m[r * n + c], v[c]));
    t[r] = acc;
  }
  memcpy(out, t, sizeof(uint64_t) * n);
}

/* Gauss-Jordan inverse mod p of an n x n matrix (n <= 12); returns 0 if
 * singular. */
static int mat_inv(const uint64_t *m, int n, uint64_t *out) {
  uint64_t a[PO_T][2 * PO_T];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 2 * n; ++j) a[i][j] = j < n ? m[i * n + j] % PO_P : (uint64_t)(j - n == i);
  for (int col = 0; col < n; ++col) {
    int piv = -1;
    for (int r = col; r < n; ++r)
      if (a[r][col]) {
        piv = r;
        break;
      }
    if (piv < 0) return 0;
    for (int j = 0; j < 2 * n; ++j) {
      uint64_t t = a[col][j];
      a[col][j] = a[piv][j];
      a[piv][j] = t;
    }
    const uint64_t iv = po_pow(a[col][col], PO_P - 2);
    for (int j = 0; j < 2 * n; ++j) a[col][j] = po_mul(a[col][j], iv);
    for (int r = 0; r < n; ++r) {
      if (r == col || !a[r][col]) continue;
      const uint64_t f = a[r][col];
      for (int j = 0; j < 2 * n; ++j) a[r][j] = po_sub(a[r][j], po_mul(f, a[col][j]));
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) out[i * n + j] = a[i][n + j];
  return 1;
}

/* Textbook permutation: every round adds its 12 constants, applies the S-box
 * to all 12 (full rounds 0-3, 26-29) or to element 0 (partial rounds 4-25),
 * then the MDS. */
void pos_permute_textbook(uint64_t *s, const uint64_t *rc) {
  uint64_t m[PO_T * PO_T];
  for (int r = 0; r < PO_T; ++r)
    for (int c = 0; c < PO_T; ++c) m[r * PO_T + c] = mds_entry(r, c);
  for (int round = 0; round < PO_ROUNDS; ++round) {
    const int full = round < PO_HALF_FULL || round >= PO_HALF_FULL + PO_PARTIAL;
    for (int i = 0; i < PO_T; ++i) {
      s[i] = po_add(s[i], rc[round * PO_T + i]);
      if (full || i == 0) s[i] = po_sbox(s[i]);
    }
    mat_vec(m, PO_T, s, s);
  }
}

/* Fused constants, in the order the tables of
 * native/poseidon_constants_template.cuh hold them. */
typedef struct {
  uint64_t rcf[PO_T];                      /* ROUND_CONSTANTS_FUSED_LAST_FULL_AND_FIRST_PARTIAL */
  uint64_t dense[PO_T * PO_T];             /* FUSED_DENSE_MATRIX_LAST_FULL_AND_FIRST_PARTIAL */
  uint64_t sbox_c[PO_PARTIAL];             /* ROUND_CONSTANTS_FOR_FUSED_S_BOXES */
  uint64_t vs[PO_PARTIAL * (PO_T - 1)];    /* VS_FOR_PARTIAL_ROUNDS */
  uint64_t w_hats[PO_PARTIAL * (PO_T - 1)]; /* W_HATS_FOR_PARTIAL_ROUNDS */
} pos_fused;

size_t pos_fused_words(void) { return sizeof(pos_fused) / sizeof(uint64_t); }

/* Derives the fused form from the 360 textbook constants.  out holds
 * pos_fused_words() words in the layout above.  Returns 1 on success. */
int pos_derive(const uint64_t *rc, uint64_t *out) {
  pos_fused f;
  uint64_t m[PO_T * PO_T], minv[PO_T * PO_T];
  for (int r = 0; r < PO_T; ++r)
    for (int c = 0; c < PO_T; ++c) m[r * PO_T + c] = mds_entry(r, c);
  if (!mat_inv(m, PO_T, minv)) return 0;
  /* constants: c_26 (the first constants of the second full half) and each
   * partial round's constants move back through M^-1; element 0 of the moved
   * vector is added after the S-box of the round before, elements 1..11 join
   * that round's own constants (they pass its S-box, which touches element 0
   * only).  What reaches round 4 is moved once more, into round 3 after its
   * S-box. */
  uint64_t cur[PO_T], e[PO_T];
  memcpy(cur, rc + (PO_HALF_FULL + PO_PARTIAL) * PO_T, sizeof cur);
  for (int k = PO_HALF_FULL + PO_PARTIAL - 1; k >= PO_HALF_FULL; --k) {
    mat_vec(minv, PO_T, cur, e);
    f.sbox_c[k - PO_HALF_FULL] = e[0];
    for (int i = 0; i < PO_T; ++i) cur[i] = i ? po_add(rc[k * PO_T + i], e[i]) : rc[k * PO_T + i] % PO_P;
  }
  mat_vec(minv, PO_T, cur, f.rcf);
  /* matrices: N = M for the last partial round; N = [[n00, n^T], [w, N^]] =
   * B A with A = diag(1, N^) and B = [[n00, n^T N^-1], [w, I]]; A commutes
   * with the partial S-box layer, so the round before gets N = A M, and round
   * 3 (full) gets the dense A_4 M.  n00 = M00 = 1 in every N. */
  uint64_t n[PO_T * PO_T];
  memcpy(n, m, sizeof n);
  for (int k = PO_HALF_FULL + PO_PARTIAL - 1; k >= PO_HALF_FULL; --k) {
    const int pr = k - PO_HALF_FULL;
    uint64_t nh[(PO_T - 1) * (PO_T - 1)], nhinv[(PO_T - 1) * (PO_T - 1)];
    for (int i = 0; i < PO_T - 1; ++i)
      for (int j = 0; j < PO_T - 1; ++j) nh[i * (PO_T - 1) + j] = n[(i + 1) * PO_T + (j + 1)];
    if (!mat_inv(nh, PO_T - 1, nhinv)) return 0;
    for (int j = 0; j < PO_T - 1; ++j) {
      uint64_t acc = 0;
      for (int i = 0; i < PO_T - 1; ++i) acc = po_add(acc, po_mul(n[i + 1], nhinv[i * (PO_T - 1) + j]));
      f.vs[pr * (PO_T - 1) + j] = acc;
      f.w_hats[pr * (PO_T - 1) + j] = n[(j + 1) * PO_T];
    }
    uint64_t next[PO_T * PO_T];
    for (int r = 0; r < PO_T; ++r)
      for (int c = 0; c < PO_T; ++c) {
        if (r == 0) {
          next[c] = m[c];
          continue;
        }
        uint64_t acc = 0;
        for (int i = 1; i < PO_T; ++i) acc = po_add(acc, po_mul(nh[(r - 1) * (PO_T - 1) + (i - 1)], m[i * PO_T + c]));
        next[r * PO_T + c] = acc;
      }
    memcpy(n, next, sizeof n);
  }
  memcpy(f.dense, n, sizeof n);
  memcpy(out, &f, sizeof f);
  return 1;
}

/* The permutation as native/poseidon_single_thread.cu:14-32 runs it, from
 * the textbook constants of rounds 0-3 and 27-29 (rc) and the fused form
 * (pos_derive's layout). */
void pos_permute_fused(uint64_t *s, const uint64_t *rc, const uint64_t *fused) {
  pos_fused f;
  memcpy(&f, fused, sizeof f);
  uint64_t m[PO_T * PO_T];
  for (int r = 0; r < PO_T; ++r)
    for (int c = 0; c < PO_T; ++c) m[r * PO_T + c] = mds_entry(r, c);
  for (int round = 0; round < PO_ROUNDS; ++round) {
    if (round < PO_HALF_FULL || round >= PO_HALF_FULL + PO_PARTIAL) {
      for (int i = 0; i < PO_T; ++i) {
        if (round != PO_HALF_FULL + PO_PARTIAL) s[i] = po_add(s[i], rc[round * PO_T + i]);
        s[i] = po_sbox(s[i]);
      }
      if (round == PO_HALF_FULL - 1) {
        for (int i = 0; i < PO_T; ++i) s[i] = po_add(s[i], f.rcf[i]);
        mat_vec(f.dense, PO_T, s, s);
      } else {
        mat_vec(m, PO_T, s, s);
      }
    } else {
      const int pr = round - PO_HALF_FULL;
      uint64_t v[PO_T];
      s[0] = po_add(po_sbox(s[0]), f.sbox_c[pr]);
      memcpy(v, s, sizeof v);
      uint64_t acc = v[0];
      for (int i = 0; i < PO_T - 1; ++i) acc = po_add(acc, po_mul(v[i + 1], f.vs[pr * (PO_T - 1) + i]));
      s[0] = acc;
      for (int i = 1; i < PO_T; ++i) s[i] = po_add(po_mul(v[0], f.w_hats[pr * (PO_T - 1) + i - 1]), v[i]);
    }
  }
}

/* Leaf hashes (native/poseidon_utils.cuh:8-42): leaf g absorbs, in chunks of
 * 8 (state[0..8] overwritten, capacity kept, a permutation after each
 * chunk, the last chunk zero-padded), the values (row r, column c) at
 * values[g * rows + r + c * rows * count] in the order offset = c * rows + r.
 * Output: state[0..4] (or the capacity state[8..12] when store_intermediate)
 * at results[g + i * count]; load_intermediate starts from the capacity at
 * results[g + i * count]. */
void pos_leaves(const uint64_t *values, uint64_t *results, uint32_t rows, uint32_t cols, uint32_t count,
                int load_intermediate, int store_intermediate, const uint64_t *rc, const uint64_t *fused) {
  for (uint32_t g = 0; g < count; ++g) {
    uint64_t s[PO_T] = {0};
    if (load_intermediate)
      for (int i = 0; i < PO_CAP; ++i) s[PO_RATE + i] = results[g + (uint64_t)i * count];
    const uint64_t total = (uint64_t)rows * cols;
    for (uint64_t off = 0; off < total;) {
      for (int i = 0; i < PO_RATE; ++i, ++off) {
        const uint64_t r = off % rows, c = off / rows;
        s[i] = c < cols ? values[(uint64_t)g * rows + r + c * (uint64_t)rows * count] : 0;
      }
      pos_permute_fused(s, rc, fused);
    }
    for (int i = 0; i < PO_CAP; ++i)
      results[g + (uint64_t)i * count] = store_intermediate ? s[PO_RATE + i] : s[i];
  }
}

/* Node hashes (native/poseidon_utils.cuh:44-61): node g hashes children 2g
 * and 2g + 1 of the layer below (4 columns of 2 * count digests): state[i] =
 * left[i], state[4 + i] = right[i], capacity zero; output state[0..4] at
 * results[g + i * count]. */
void pos_nodes(const uint64_t *values, uint64_t *results, uint32_t count, const uint64_t *rc, const uint64_t *fused) {
  for (uint32_t g = 0; g < count; ++g) {
    uint64_t s[PO_T] = {0};
    for (int i = 0; i < PO_CAP; ++i)
      for (int j = 0; j < 2; ++j) s[j * PO_CAP + i] = values[2 * ((uint64_t)g + (uint64_t)i * count) + j];
    pos_permute_fused(s, rc, fused);
    for (int i = 0; i < PO_CAP; ++i) results[g + (uint64_t)i * count] = s[i];
  }
}

/* 360 constants by the Poseidon reference parameter procedure: an 80-bit
 * Grain LFSR seeded with (field = 1, S-box = 0, field bits = 64, width = 12,
 * full rounds = 8, partial rounds = 22, thirty 1-bits), 160 bits discarded,
 * then self-shrinking (bit pairs (a, b): b is output when a = 1), 64-bit
 * big-endian draws, values >= p rejected. */
void pos_grain_constants(uint64_t *out) {
  uint8_t s[80];
  int k = 0;
  const struct {
    unsigned v, w;
  } fields[6] = {{1, 2}, {0, 4}, {64, 12}, {12, 12}, {8, 10}, {22, 10}};
  for (int f = 0; f < 6; ++f)
    for (int b = (int)fields[f].w - 1; b >= 0; --b) s[k++] = (fields[f].v >> b) & 1u;
  while (k < 80) s[k++] = 1;
  int head = 0; /* s[head] is the oldest bit */
#define GRAIN_BIT(i) s[(head + (i)) % 80]
#define GRAIN_STEP(dst)                                                                                  \
  do {                                                                                                   \
    const uint8_t nb = GRAIN_BIT(62) ^ GRAIN_BIT(51) ^ GRAIN_BIT(38) ^ GRAIN_BIT(23) ^ GRAIN_BIT(13) ^ \
                       GRAIN_BIT(0);                                                                     \
    s[head] = nb;                                                                                        \
    head = (head + 1) % 80;                                                                              \
    dst = nb;                                                                                            \
  } while (0)
  uint8_t tmp;
  for (int i = 0; i < 160; ++i) GRAIN_STEP(tmp);
  int produced = 0;
  while (produced < PO_ROUNDS * PO_T) {
    uint64_t v = 0;
    for (int b = 0; b < 64; ++b) {
      uint8_t a, bit;
      for (;;) {
        GRAIN_STEP(a);
        GRAIN_STEP(bit);
        if (a) break;
      }
      v = (v << 1) | bit;
    }
    if (v < PO_P) out[produced++] = v;
  }
#undef GRAIN_STEP
#undef GRAIN_BIT
  (void)tmp;
}

/* ---- Poseidon2 (native/poseidon2_single_thread.cu:7-24,
 * native/poseidon2_single_thread.cuh:12-66) ----
 * Same width, rounds, S-box and round constants (the reference's
 * poseidon_common::ALL_ROUND_CONSTANTS serves both hashes).  Textbook form
 * with explicit matrices:
 *   - external M_E = circ(2 M4, M4, M4) over the three 4-tiles, M4 =
 *     [[5,7,1,3],[4,6,1,1],[1,3,5,7],[1,1,4,6]] (eprint 2023/323, App. B);
 *   - internal M_I = J + diag(2^mu_i), mu = LOG_MU_MINUS_ONE
 *     (poseidon_constants_template.cuh:125), so M_I[i][i] = 2^mu_i + 1;
 *   - M_E first, then full rounds (constants on all 12, S-box on all, M_E)
 *     0-3 and 26-29, partial rounds 4-25 (constant and S-box on element 0,
 *     M_I). */
static const uint64_t P2_M4[4][4] = {{5, 7, 1, 3}, {4, 6, 1, 1}, {1, 3, 5, 7}, {1, 1, 4, 6}};
static const unsigned P2_MU[PO_T] = {4, 14, 11, 8, 0, 5, 2, 9, 13, 6, 3, 12};

void pos2_matrices(uint64_t *m_e, uint64_t *m_i) {
  for (int r = 0; r < PO_T; ++r)
    for (int c = 0; c < PO_T; ++c) {
      const uint64_t m4 = P2_M4[r % 4][c % 4];
      m_e[r * PO_T + c] = (r / 4 == c / 4) ? 2 * m4 : m4;
      m_i[r * PO_T + c] = 1 + (r == c ? (1ull << P2_MU[r]) : 0);
    }
}

void pos2_permute(uint64_t *s, const uint64_t *rc) {
  uint64_t m_e[PO_T * PO_T], m_i[PO_T * PO_T];
  pos2_matrices(m_e, m_i);
  for (int i = 0; i < PO_T; ++i) s[i] %= PO_P;
  mat_vec(m_e, PO_T, s, s);
  for (int round = 0; round < PO_ROUNDS; ++round) {
    const int full = round < PO_HALF_FULL || round >= PO_HALF_FULL + PO_PARTIAL;
    for (int i = 0; i < PO_T; ++i) {
      if (full || i == 0) s[i] = po_sbox(po_add(s[i], rc[round * PO_T + i]));
    }
    mat_vec(full ? m_e : m_i, PO_T, s, s);
  }
}

/* Leaves and nodes as pos_leaves / pos_nodes, with the Poseidon2
 * permutation. */
void pos2_leaves(const uint64_t *values, uint64_t *results, uint32_t rows, uint32_t cols, uint32_t count,
                 int load_intermediate, int store_intermediate, const uint64_t *rc) {
  for (uint32_t g = 0; g < count; ++g) {
    uint64_t s[PO_T] = {0};
    if (load_intermediate)
      for (int i = 0; i < PO_CAP; ++i) s[PO_RATE + i] = results[g + (uint64_t)i * count];
    const uint64_t total = (uint64_t)rows * cols;
    for (uint64_t off = 0; off < total;) {
      for (int i = 0; i < PO_RATE; ++i, ++off) {
        const uint64_t r = off % rows, c = off / rows;
        s[i] = c < cols ? values[(uint64_t)g * rows + r + c * (uint64_t)rows * count] : 0;
      }
      pos2_permute(s, rc);
    }
    for (int i = 0; i < PO_CAP; ++i)
      results[g + (uint64_t)i * count] = store_intermediate ? s[PO_RATE + i] : s[i];
  }
}

void pos2_nodes(const uint64_t *values, uint64_t *results, uint32_t count, const uint64_t *rc) {
  for (uint32_t g = 0; g < count; ++g) {
    uint64_t s[PO_T] = {0};
    for (int i = 0; i < PO_CAP; ++i)
      for (int j = 0; j < 2; ++j) s[j * PO_CAP + i] = values[2 * ((uint64_t)g + (uint64_t)i * count) + j];
    pos2_permute(s, rc);
    for (int i = 0; i < PO_CAP; ++i) results[g + (uint64_t)i * count] = s[i];
  }
}

/*
 * pgw_oracle.c -- CPU restatement (plain C) of the reference's sequence path.
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker / CPU baseline.  The product
 * library never links or calls it.  See pgw_oracle.h for what pins each part.
 *
 * Citations are to wilthoma/prym_green_ancillary.  The structure deliberately
 * follows the reference's single-threaded CPU oracle
 * (cuda/cuprym_cyclic.cu:1402-1966): row gathers, u64 sums, one reduction per
 * output.  Generalised to p < 2^31 (u128 partial sums where needed).
 */
#include "pgw_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* combinatorics                                                            */
/* ------------------------------------------------------------------------ */

/* cuda/cuprym_cyclic.cu:134-147 */
uint64_t orc_binom(uint32_t n, uint32_t k) {
    if (k > n) return 0;
    if (k > n - k) k = n - k;
    u128 out = 1;
    for (uint32_t i = 0; i < k; ++i) out = (out * (n - i)) / (i + 1);
    return (uint64_t)out;
}

/* Lexicographic unrank, cuda/cuprym_cyclic.cu:184-211 (src/subsets.rs:18-61). */
int orc_unrank_subset(uint64_t rank, uint32_t n, uint32_t k, uint32_t* out) {
    if (rank >= orc_binom(n, k)) return -1;
    uint32_t start = 0;
    for (uint32_t pos = 0; pos < k; ++pos) {
        uint32_t remaining = k - pos - 1;
        int chosen = 0;
        for (uint32_t value = start; value < n; ++value) {
            uint64_t count = orc_binom(n - value - 1, remaining);
            if (rank < count) {
                out[pos] = value;
                start = value + 1;
                chosen = 1;
                break;
            }
            rank -= count;
        }
        if (!chosen) return -1;
    }
    return 0;
}

/* Subset masks of all k-subsets of [0,n) in lex order, plus a mask-sorted
 * index for rank lookups (the reference uses an unordered_map keyed by mask,
 * cuda/cuprym_cyclic.cu:653-668). */
typedef struct {
    uint64_t mask;
    uint32_t rank;
} mask_rank;

static int cmp_mask_rank(const void* a, const void* b) {
    uint64_t x = ((const mask_rank*)a)->mask, y = ((const mask_rank*)b)->mask;
    return x < y ? -1 : (x > y ? 1 : 0);
}

typedef struct {
    uint32_t n, k;
    size_t count;
    uint64_t* masks;     /* by rank */
    mask_rank* sorted;   /* by mask */
} subset_table;

static void subset_table_build(subset_table* t, uint32_t n, uint32_t k) {
    t->n = n;
    t->k = k;
    t->count = (size_t)orc_binom(n, k);
    t->masks = (uint64_t*)malloc(t->count * sizeof(uint64_t) + 8);
    t->sorted = (mask_rank*)malloc(t->count * sizeof(mask_rank) + 8);
    uint32_t s[64];
    for (size_t rank = 0; rank < t->count; ++rank) {
        orc_unrank_subset(rank, n, k, s);
        uint64_t mask = 0;
        for (uint32_t j = 0; j < k; ++j) mask |= 1ULL << s[j];
        t->masks[rank] = mask;
        t->sorted[rank].mask = mask;
        t->sorted[rank].rank = (uint32_t)rank;
    }
    qsort(t->sorted, t->count, sizeof(mask_rank), cmp_mask_rank);
}

static uint32_t subset_table_rank(const subset_table* t, uint64_t mask) {
    size_t lo = 0, hi = t->count;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (t->sorted[mid].mask < mask) lo = mid + 1; else hi = mid;
    }
    if (lo >= t->count || t->sorted[lo].mask != mask) {
        fprintf(stderr, "oracle: subset rank lookup failed\n");
        abort();
    }
    return t->sorted[lo].rank;
}

static void subset_table_free(subset_table* t) {
    free(t->masks);
    free(t->sorted);
}

/* ------------------------------------------------------------------------ */
/* RNG and hashes                                                           */
/* ------------------------------------------------------------------------ */

/* cuda/cuprym_cyclic.cu:1153-1158 (Vigna's splitmix64) */
uint64_t orc_splitmix64_next(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* cuda/cuprym_cyclic.cu:1143-1151 */
uint64_t orc_seeded_sector_value(uint64_t seed, uint64_t sector) {
    uint64_t x = seed ^ (0x9E3779B97F4A7C15ULL + (sector << 6) + (sector >> 2));
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBULL;
    x ^= x >> 31;
    return x;
}

/* cuda/cuprym_cyclic.cu:1207-1218 */
uint64_t orc_rowmix_bounded(uint64_t* state, uint64_t bound) {
    uint64_t threshold = (0ULL - bound) % bound;
    for (;;) {
        uint64_t value = orc_splitmix64_next(state);
        if (value >= threshold) return value % bound;
    }
}

#define FNV_OFFSET 14695981039346656037ULL
#define FNV_PRIME 1099511628211ULL

static uint64_t fnv_u8(uint64_t h, uint8_t v) { h ^= v; return h * FNV_PRIME; }
static uint64_t fnv_u64(uint64_t h, uint64_t v) {
    for (int byte = 0; byte < 8; ++byte) h = fnv_u8(h, (uint8_t)((v >> (8 * byte)) & 0xff));
    return h;
}
static uint64_t fnv_str(uint64_t h, const char* s) {
    for (; *s; ++s) h = fnv_u8(h, (uint8_t)*s);
    return h;
}

uint64_t orc_fnv1a64_string(const char* s) { return fnv_str(FNV_OFFSET, s); }

/* cuda/cuprym_cyclic.cu:1181-1188 */
uint64_t orc_fnv1a64_u32_vector(const uint32_t* values, size_t count) {
    uint64_t h = fnv_u64(FNV_OFFSET, (uint64_t)count);
    for (size_t i = 0; i < count; ++i) h = fnv_u64(h, (uint64_t)values[i]);
    return h;
}

/* mt19937_64 (Matsumoto & Nishimura 2000; std::mt19937_64 parameters). */
typedef struct {
    uint64_t mt[312];
    int idx;
} mt64;

static void mt64_seed(mt64* g, uint64_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 312; ++i)
        g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->idx = 312;
}

static uint64_t mt64_next(mt64* g) {
    const uint64_t upper = 0xFFFFFFFF80000000ULL, lower = 0x7FFFFFFFULL;
    if (g->idx >= 312) {
        for (int i = 0; i < 312; ++i) {
            uint64_t x = (g->mt[i] & upper) | (g->mt[(i + 1) % 312] & lower);
            uint64_t xa = x >> 1;
            if (x & 1ULL) xa ^= 0xB5026F5AA96619E9ULL;
            g->mt[i] = g->mt[(i + 156) % 312] ^ xa;
        }
        g->idx = 0;
    }
    uint64_t y = g->mt[g->idx++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= y >> 43;
    return y;
}

void orc_mt19937_64_outputs(uint64_t seed, size_t count, uint64_t* out) {
    mt64 g;
    mt64_seed(&g, seed);
    for (size_t i = 0; i < count; ++i) out[i] = mt64_next(&g);
}

/* std::uniform_int_distribution<uint64_t>(lo, hi) over a 64-bit engine as
 * libstdc++ 11 implements it: Lemire's nearly-divisionless multiply-shift
 * ("Fast Random Integer Generation in an Interval", ACM TOMACS 2019) with a
 * 128-bit product.  Used by generate_random_vector (cuda/matrices.h:44-56). */
static uint64_t uniform_draw(mt64* g, uint64_t lo, uint64_t hi) {
    uint64_t s = hi - lo + 1;
    if (s == 0) return mt64_next(g); /* full 64-bit range */
    u128 product = (u128)mt64_next(g) * s;
    uint64_t low = (uint64_t)product;
    if (low < s) {
        uint64_t threshold = (0ULL - s) % s;
        while (low < threshold) {
            product = (u128)mt64_next(g) * s;
            low = (uint64_t)product;
        }
    }
    return (uint64_t)(product >> 64) + lo;
}

void orc_uniform_u64(uint64_t seed, uint64_t lo, uint64_t hi, size_t count, uint64_t* out) {
    mt64 g;
    mt64_seed(&g, seed);
    for (size_t i = 0; i < count; ++i) out[i] = uniform_draw(&g, lo, hi);
}

/* cuda/cuprym_cyclic.cu:2196-2199 */
void orc_reference_random_state(uint64_t seed, uint64_t sector, uint32_t p,
                                size_t rows, size_t cols, size_t b,
                                uint32_t* dr, uint32_t* dc, uint32_t* v) {
    mt64 g;
    mt64_seed(&g, orc_seeded_sector_value(seed, sector));
    for (size_t i = 0; i < rows; ++i) dr[i] = (uint32_t)uniform_draw(&g, 1, p - 1);
    for (size_t i = 0; i < cols; ++i) dc[i] = (uint32_t)uniform_draw(&g, 1, p - 1);
    for (size_t i = 0; i < cols * b; ++i) v[i] = (uint32_t)uniform_draw(&g, 0, p - 1);
}

/* ------------------------------------------------------------------------ */
/* row mixing                                                               */
/* ------------------------------------------------------------------------ */

/* cuda/cuprym_cyclic.cu:1196-1205 */
uint64_t orc_rowmix_realized_seed(uint64_t seed, const char* op, uint64_t sector,
                                  uint64_t rows, uint32_t p) {
    uint64_t h = FNV_OFFSET;
    h = fnv_str(h, "paired-sl2-target-v1");
    h = fnv_str(h, op);
    h = fnv_u64(h, seed);
    h = fnv_u64(h, sector);
    h = fnv_u64(h, rows);
    h = fnv_u64(h, (uint64_t)p);
    return orc_seeded_sector_value(h, sector);
}

/* cuda/cuprym_cyclic.cu:1227-1268 (a, b widened to u32 for p > 65535) */
void orc_rowmix_plan(uint64_t realized_seed, size_t rows, size_t rounds, uint32_t p,
                     uint32_t* pairs) {
    uint32_t* perm = (uint32_t*)malloc((rows + 1) * sizeof(uint32_t));
    for (size_t i = 0; i < rows; ++i) perm[i] = (uint32_t)i;
    uint64_t state = realized_seed;
    size_t pair_count = rows / 2;
    for (size_t round = 0; round < rounds; ++round) {
        for (size_t i = rows; i > 1; --i) {
            size_t j = (size_t)orc_rowmix_bounded(&state, i);
            uint32_t tmp = perm[i - 1];
            perm[i - 1] = perm[j];
            perm[j] = tmp;
        }
        uint32_t* out = pairs + round * pair_count * 4;
        for (size_t q = 0; q < pair_count; ++q) {
            out[4 * q + 0] = perm[2 * q];
            out[4 * q + 1] = perm[2 * q + 1];
            out[4 * q + 2] = (uint32_t)(1 + orc_rowmix_bounded(&state, p - 1));
            out[4 * q + 3] = (uint32_t)(1 + orc_rowmix_bounded(&state, p - 1));
        }
    }
    free(perm);
}

/* cuda/cuprym_cyclic.cu:1270-1287 */
uint64_t orc_rowmix_plan_hash(size_t rows, size_t rounds, uint64_t seed,
                              uint64_t realized_seed, const uint32_t* pairs) {
    uint64_t h = FNV_OFFSET;
    h = fnv_str(h, "paired-sl2-target-v1");
    h = fnv_u64(h, rows);
    h = fnv_u64(h, rounds);
    h = fnv_u64(h, seed);
    h = fnv_u64(h, realized_seed);
    size_t pair_count = rows / 2;
    for (size_t round = 0; round < rounds; ++round) {
        h = fnv_u64(h, pair_count);
        const uint32_t* pr = pairs + round * pair_count * 4;
        for (size_t q = 0; q < pair_count; ++q)
            for (int f = 0; f < 4; ++f) h = fnv_u64(h, pr[4 * q + f]);
    }
    return h;
}

void orc_synthetic_table(uint64_t seed, uint32_t p, size_t count, uint32_t* out) {
    uint64_t state = seed;
    for (size_t i = 0; i < count; ++i) out[i] = (uint32_t)orc_rowmix_bounded(&state, p);
}

/* ------------------------------------------------------------------------ */
/* the matrix-free Koszul operator (one "phi" differential)                 */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint32_t i;      /* original V index */
    uint32_t other;  /* rank of the neighbouring subset */
    uint8_t negative;
} incidence;

/* One differential Lambda^d U (x) A0 -> Lambda^(d-1) U (x) A1, restricted to a
 * sector, as CyclicPhiHostData (cyclic_cuda_helpers.h:46-75) built by
 * build_common_incidences_generic + build_sector_layout_generic
 * (cuda/cuprym_cyclic.cu:631-819). */
typedef struct {
    uint32_t nI, nB, nA, r;       /* table dims mu[nI][nB][nA] */
    const uint32_t* mu;
    const uint16_t* wV;           /* by original index */
    const uint16_t* wA0;
    const uint16_t* wA1;
    /* character maps */
    uint32_t* a0_off; uint16_t* a0_idx;
    uint32_t* a1_off; uint16_t* a1_idx;
    /* subsets */
    size_t row_subsets, col_subsets, row_width, col_width;
    uint16_t* row_subset_w;
    uint16_t* col_subset_w;
    incidence* row_inc;   /* row_subsets * row_width */
    incidence* col_inc;   /* col_subsets * col_width */
    /* sector layout */
    size_t rows, cols;
    uint32_t* row_off;  uint32_t* row_rank;  uint16_t* row_beta;
    uint32_t* col_off;  uint32_t* col_rank;  uint16_t* col_a;
    const uint32_t* row_scale;  /* NULL = ones */
} phi_op;

static void char_map(const uint16_t* w, uint32_t len, uint32_t r, uint32_t** off, uint16_t** idx) {
    *off = (uint32_t*)malloc((r + 1) * sizeof(uint32_t));
    *idx = (uint16_t*)malloc((len + 1) * sizeof(uint16_t));
    uint32_t count = 0;
    for (uint32_t ch = 0; ch < r; ++ch) {
        (*off)[ch] = count;
        for (uint32_t j = 0; j < len; ++j)
            if (w[j] == ch) (*idx)[count++] = (uint16_t)j;
    }
    (*off)[r] = count;
}

static uint32_t sub_mod_u(uint32_t a, uint32_t b, uint32_t r) { return (a + r - (b % r)) % r; }

/* local subsets over `local_n` elements mapped through original[], source
 * degree d (columns), target degree d-1 (rows). */
static void phi_build(phi_op* op, uint32_t nI, uint32_t nB, uint32_t nA, uint32_t r,
                      const uint32_t* mu, const uint16_t* wV, const uint16_t* wA0,
                      const uint16_t* wA1, uint32_t local_n, const uint32_t* original,
                      uint32_t d, uint32_t sector) {
    memset(op, 0, sizeof(*op));
    op->nI = nI; op->nB = nB; op->nA = nA; op->r = r;
    op->mu = mu; op->wV = wV; op->wA0 = wA0; op->wA1 = wA1;
    char_map(wA0, nA, r, &op->a0_off, &op->a0_idx);
    char_map(wA1, nB, r, &op->a1_off, &op->a1_idx);

    subset_table rows_t, cols_t;
    subset_table_build(&rows_t, local_n, d - 1);
    subset_table_build(&cols_t, local_n, d);
    op->row_subsets = rows_t.count;
    op->col_subsets = cols_t.count;
    op->row_width = local_n - (d - 1);
    op->col_width = d;
    op->row_subset_w = (uint16_t*)malloc(rows_t.count * sizeof(uint16_t) + 2);
    op->col_subset_w = (uint16_t*)malloc(cols_t.count * sizeof(uint16_t) + 2);
    for (size_t s = 0; s < rows_t.count; ++s) {
        uint32_t sum = 0;
        for (uint32_t j = 0; j < local_n; ++j)
            if (rows_t.masks[s] >> j & 1ULL) sum += wV[original[j]];
        op->row_subset_w[s] = (uint16_t)(sum % r);
    }
    for (size_t s = 0; s < cols_t.count; ++s) {
        uint32_t sum = 0;
        for (uint32_t j = 0; j < local_n; ++j)
            if (cols_t.masks[s] >> j & 1ULL) sum += wV[original[j]];
        op->col_subset_w[s] = (uint16_t)(sum % r);
    }
    /* row incidences: for each target subset, local_i not in it, in
     * increasing local_i; sign = parity of #elements below local_i
     * (cuda/cuprym_cyclic.cu:671-699) */
    op->row_inc = (incidence*)malloc(rows_t.count * op->row_width * sizeof(incidence) + 16);
    size_t q = 0;
    for (size_t s = 0; s < rows_t.count; ++s) {
        uint64_t mask = rows_t.masks[s];
        for (uint32_t li = 0; li < local_n; ++li) {
            if (mask >> li & 1ULL) continue;
            uint32_t position = (uint32_t)__builtin_popcountll(mask & ((1ULL << li) - 1));
            op->row_inc[q].i = original[li];
            op->row_inc[q].other = subset_table_rank(&cols_t, mask | (1ULL << li));
            op->row_inc[q].negative = (uint8_t)(position & 1);
            ++q;
        }
    }
    /* column incidences (cuda/cuprym_cyclic.cu:701-717) */
    op->col_inc = (incidence*)malloc(cols_t.count * op->col_width * sizeof(incidence) + 16);
    q = 0;
    for (size_t s = 0; s < cols_t.count; ++s) {
        uint64_t mask = cols_t.masks[s];
        uint32_t position = 0;
        for (uint32_t li = 0; li < local_n; ++li) {
            if (!(mask >> li & 1ULL)) continue;
            op->col_inc[q].i = original[li];
            op->col_inc[q].other = subset_table_rank(&rows_t, mask & ~(1ULL << li));
            op->col_inc[q].negative = (uint8_t)(position & 1);
            ++q;
            ++position;
        }
    }
    subset_table_free(&rows_t);
    subset_table_free(&cols_t);

    /* sector layout (cuda/cuprym_cyclic.cu:733-790) */
    size_t nrows = 0, ncols = 0;
    for (size_t s = 0; s < op->row_subsets; ++s) {
        uint32_t ch = sub_mod_u(sector, op->row_subset_w[s], r);
        nrows += op->a1_off[ch + 1] - op->a1_off[ch];
    }
    for (size_t s = 0; s < op->col_subsets; ++s) {
        uint32_t ch = sub_mod_u(sector, op->col_subset_w[s], r);
        ncols += op->a0_off[ch + 1] - op->a0_off[ch];
    }
    op->rows = nrows;
    op->cols = ncols;
    op->row_off = (uint32_t*)malloc((op->row_subsets + 1) * sizeof(uint32_t));
    op->col_off = (uint32_t*)malloc((op->col_subsets + 1) * sizeof(uint32_t));
    op->row_rank = (uint32_t*)malloc((nrows + 1) * sizeof(uint32_t));
    op->row_beta = (uint16_t*)malloc((nrows + 1) * sizeof(uint16_t));
    op->col_rank = (uint32_t*)malloc((ncols + 1) * sizeof(uint32_t));
    op->col_a = (uint16_t*)malloc((ncols + 1) * sizeof(uint16_t));
    size_t at = 0;
    for (size_t s = 0; s < op->row_subsets; ++s) {
        op->row_off[s] = (uint32_t)at;
        uint32_t ch = sub_mod_u(sector, op->row_subset_w[s], r);
        for (uint32_t idx = op->a1_off[ch]; idx < op->a1_off[ch + 1]; ++idx) {
            op->row_rank[at] = (uint32_t)s;
            op->row_beta[at] = op->a1_idx[idx];
            ++at;
        }
    }
    op->row_off[op->row_subsets] = (uint32_t)at;
    at = 0;
    for (size_t s = 0; s < op->col_subsets; ++s) {
        op->col_off[s] = (uint32_t)at;
        uint32_t ch = sub_mod_u(sector, op->col_subset_w[s], r);
        for (uint32_t idx = op->a0_off[ch]; idx < op->a0_off[ch + 1]; ++idx) {
            op->col_rank[at] = (uint32_t)s;
            op->col_a[at] = op->a0_idx[idx];
            ++at;
        }
    }
    op->col_off[op->col_subsets] = (uint32_t)at;
}

static void phi_free(phi_op* op) {
    free(op->a0_off); free(op->a0_idx); free(op->a1_off); free(op->a1_idx);
    free(op->row_subset_w); free(op->col_subset_w);
    free(op->row_inc); free(op->col_inc);
    free(op->row_off); free(op->col_off);
    free(op->row_rank); free(op->row_beta); free(op->col_rank); free(op->col_a);
}

static inline uint32_t mulmod(uint32_t a, uint32_t b, uint32_t p) {
    return (uint32_t)(((uint64_t)a * b) % p);
}

/* cpu_phi_forward_scaled_columns, cuda/cuprym_cyclic.cu:1591-1637, restricted
 * to output rows [r0, r1). */
static void phi_forward(const phi_op* op, const uint32_t* in, uint32_t* out, size_t b,
                        uint32_t p, size_t r0, size_t r1) {
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t row = r0; row < r1; ++row) {
        uint32_t beta = op->row_beta[row];
        uint32_t bw = op->wA1[beta];
        const incidence* inc = op->row_inc + (size_t)op->row_rank[row] * op->row_width;
        for (size_t c = 0; c < b; ++c) {
            u128 sum = 0;
            for (size_t pos = 0; pos < op->row_width; ++pos) {
                uint32_t q = sub_mod_u(bw, op->wV[inc[pos].i], op->r);
                uint32_t start = op->a0_off[q], end = op->a0_off[q + 1];
                uint32_t base = op->col_off[inc[pos].other];
                for (uint32_t idx = start; idx < end; ++idx) {
                    uint32_t a = op->a0_idx[idx];
                    uint32_t raw = op->mu[((size_t)inc[pos].i * op->nB + beta) * op->nA + a];
                    if (raw == 0) continue;
                    uint32_t x = in[((size_t)base + (idx - start)) * b + c];
                    uint32_t coeff = inc[pos].negative ? p - raw : raw;
                    sum += (uint64_t)coeff * x;
                }
            }
            uint32_t red = (uint32_t)(sum % p);
            out[row * b + c] = op->row_scale ? mulmod(red, op->row_scale[row], p) : red;
        }
    }
}

/* cpu_phi_transpose, cuda/cuprym_cyclic.cu:1639-1685 (column scaling = ones). */
static void phi_transpose_range(const phi_op* op, const uint32_t* in, uint32_t* out,
                                size_t b, uint32_t p, size_t c0, size_t c1) {
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t col = c0; col < c1; ++col) {
        uint32_t a = op->col_a[col];
        uint32_t aw = op->wA0[a];
        const incidence* inc = op->col_inc + (size_t)op->col_rank[col] * op->col_width;
        for (size_t c = 0; c < b; ++c) {
            u128 sum = 0;
            for (size_t pos = 0; pos < op->col_width; ++pos) {
                uint32_t q = (aw + op->wV[inc[pos].i]) % op->r;
                uint32_t start = op->a1_off[q], end = op->a1_off[q + 1];
                uint32_t base = op->row_off[inc[pos].other];
                for (uint32_t idx = start; idx < end; ++idx) {
                    uint32_t beta = op->a1_idx[idx];
                    uint32_t raw = op->mu[((size_t)inc[pos].i * op->nB + beta) * op->nA + a];
                    if (raw == 0) continue;
                    uint32_t y = in[((size_t)base + (idx - start)) * b + c];
                    uint32_t coeff = inc[pos].negative ? p - raw : raw;
                    sum += (uint64_t)coeff * y;
                }
            }
            out[col * b + c] = (uint32_t)(sum % p);
        }
    }
}

static void phi_transpose(const phi_op* op, const uint32_t* in, uint32_t* out, size_t b,
                          uint32_t p) {
    phi_transpose_range(op, in, out, b, p, 0, op->cols);
}

/* ------------------------------------------------------------------------ */
/* composite operators                                                      */
/* ------------------------------------------------------------------------ */

enum { ORC_SYNTHETIC = 1, ORC_CYCLIC = 2 };

struct orc_op {
    int kind;
    uint32_t p;
    size_t rows, cols;          /* F is rows x cols */
    const uint32_t* dr;
    const uint32_t* dc;
    size_t rounds;
    const uint32_t* pairs;
    /* synthetic: d_op (table T) then e_op (table E, used transposed) */
    phi_op d_op, e_op;
    uint32_t* ident;            /* identity index map */
    uint16_t* zero_w;
    /* cyclic eliminated */
    phi_op dm, d2;
    uint32_t n_v, kernel_dim;
    const uint32_t* right_inverse;  /* g x n */
    const uint32_t* kernel_basis;   /* g x kdim */
    size_t z_count;
    uint32_t* z_off; uint32_t* z_rank; uint16_t* z_kappa;
    size_t x_cols;
    /* deformation variant of the cyclic operator: 0 none, 1 augmented
     * G = [F0 | -dense], 2 replacement G = [F0 without columns drop0 < drop1 |
     * dense]; dense is target rows x dense_cols, row-major */
    int deform;
    const uint32_t* dense;
    size_t dense_cols;
    size_t drop0, drop1;
    size_t base_cols;
};

orc_op* orc_synthetic_create(uint32_t n, uint32_t k, uint32_t M, uint32_t Mp, uint32_t p,
                             const uint32_t* T, const uint32_t* E,
                             const uint32_t* dr, const uint32_t* dc,
                             size_t rowmix_rounds, const uint32_t* rowmix_pairs) {
    orc_op* op = (orc_op*)calloc(1, sizeof(orc_op));
    op->kind = ORC_SYNTHETIC;
    op->p = p;
    uint32_t big = n > M ? n : M;
    if (Mp > big) big = Mp;
    op->ident = (uint32_t*)malloc(64 * sizeof(uint32_t));
    for (uint32_t j = 0; j < 64; ++j) op->ident[j] = j;
    op->zero_w = (uint16_t*)calloc(big + 1, sizeof(uint16_t));
    /* r = 1: every weight 0, every character block is the whole basis */
    phi_build(&op->d_op, n, Mp, M, 1, T, op->zero_w, op->zero_w, op->zero_w, n, op->ident, k, 0);
    phi_build(&op->e_op, n, Mp, M, 1, E, op->zero_w, op->zero_w, op->zero_w, n, op->ident, k, 0);
    op->rows = op->d_op.cols; /* A = E o D is N x N */
    op->cols = op->d_op.cols;
    op->dr = dr;
    op->dc = dc;
    op->rounds = rowmix_rounds;
    op->pairs = rowmix_pairs;
    return op;
}

orc_op* orc_cyclic_create(uint32_t genus, uint32_t r, uint32_t p, uint32_t sector,
                          const uint32_t* mu, const uint16_t* weight_v,
                          const uint16_t* weight_a0, const uint16_t* weight_a1,
                          uint32_t w_index, uint32_t w_weight, uint32_t kernel_dim,
                          const uint16_t* kernel_weights, const uint32_t* right_inverse,
                          const uint32_t* kernel_basis) {
    orc_op* op = (orc_op*)calloc(1, sizeof(orc_op));
    op->kind = ORC_CYCLIC;
    op->p = p;
    uint32_t n = genus - 3, m = genus / 2;
    op->n_v = n;
    op->kernel_dim = kernel_dim;
    op->right_inverse = right_inverse;
    op->kernel_basis = kernel_basis;
    /* U = V \ {w} (cuda/cuprym_cyclic.cu:2616-2626) */
    op->ident = (uint32_t*)malloc(64 * sizeof(uint32_t));
    uint32_t ln = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (i != w_index) op->ident[ln++] = i;
    uint32_t y_sector = sub_mod_u(sector, w_weight, r);
    phi_build(&op->dm, n, n, genus, r, mu, weight_v, weight_a0, weight_a1, ln, op->ident, m, sector);
    phi_build(&op->d2, n, n, genus, r, mu, weight_v, weight_a0, weight_a1, ln, op->ident, m - 1, y_sector);
    /* z layout over the (m-1)-subsets with the kernel characters
     * (build_tensor_layout, cuda/cuprym_cyclic.cu:792-819) */
    uint32_t* k_off; uint16_t* k_idx;
    char_map(kernel_weights, kernel_dim, r, &k_off, &k_idx);
    size_t subsets = op->dm.row_subsets;
    op->z_off = (uint32_t*)malloc((subsets + 1) * sizeof(uint32_t));
    size_t cnt = 0;
    for (size_t s = 0; s < subsets; ++s) {
        uint32_t ch = sub_mod_u(y_sector, op->dm.row_subset_w[s], r);
        cnt += k_off[ch + 1] - k_off[ch];
    }
    op->z_count = cnt;
    op->z_rank = (uint32_t*)malloc((cnt + 1) * sizeof(uint32_t));
    op->z_kappa = (uint16_t*)malloc((cnt + 1) * sizeof(uint16_t));
    size_t at = 0;
    for (size_t s = 0; s < subsets; ++s) {
        op->z_off[s] = (uint32_t)at;
        uint32_t ch = sub_mod_u(y_sector, op->dm.row_subset_w[s], r);
        for (uint32_t idx = k_off[ch]; idx < k_off[ch + 1]; ++idx) {
            op->z_rank[at] = (uint32_t)s;
            op->z_kappa[at] = k_idx[idx];
            ++at;
        }
    }
    op->z_off[subsets] = (uint32_t)at;
    free(k_off);
    free(k_idx);
    op->x_cols = op->dm.cols;
    op->cols = op->dm.cols + op->z_count;
    op->rows = op->d2.rows;
    return op;
}

/* The augmented and replacement operators of run_deformation_eliminated_variant
 * (cuda/cuprym_deformation.cu:2554-2855) on an eliminated cyclic operator:
 * augmented G = [F0 | -dense] (forward: dense axpy with SUBTRACT, :2787-2795),
 * replacement G = [F0 restricted to the base columns other than drop0 < drop1 |
 * dense] (insert/gather complement, :2796-2855).  Source columns: base then
 * dense for augmented, kept base columns then dense for replacement. */
int orc_cyclic_deformation(orc_op* op, int variant, const uint32_t* dense, size_t dense_cols,
                           size_t drop0, size_t drop1) {
    if (op->kind != ORC_CYCLIC || op->deform || (variant != 1 && variant != 2) || dense_cols == 0)
        return -1;
    op->base_cols = op->cols;
    if (variant == 2 && !(drop0 < drop1 && drop1 < op->base_cols)) return -1;
    op->deform = variant;
    op->dense = dense;
    op->dense_cols = dense_cols;
    op->drop0 = drop0;
    op->drop1 = drop1;
    op->cols = (variant == 1 ? op->base_cols : op->base_cols - 2) + dense_cols;
    return 0;
}

void orc_set_preconditioners(orc_op* op, const uint32_t* dr, const uint32_t* dc,
                             size_t rowmix_rounds, const uint32_t* rowmix_pairs) {
    op->dr = dr;
    op->dc = dc;
    op->rounds = rowmix_rounds;
    op->pairs = rowmix_pairs;
}

void orc_op_dims(const orc_op* op, size_t* rows, size_t* cols) {
    *rows = op->rows;
    *cols = op->cols;
}

void orc_op_destroy(orc_op* op) {
    if (!op) return;
    if (op->kind == ORC_SYNTHETIC) {
        phi_free(&op->d_op);
        phi_free(&op->e_op);
        free(op->zero_w);
    } else {
        phi_free(&op->dm);
        phi_free(&op->d2);
        free(op->z_off); free(op->z_rank); free(op->z_kappa);
    }
    free(op->ident);
    free(op);
}

static void scale_rows(const uint32_t* in, const uint32_t* s, uint32_t* out, size_t rows,
                       size_t b, uint32_t p) {
    for (size_t r = 0; r < rows; ++r)
        for (size_t c = 0; c < b; ++c) out[r * b + c] = mulmod(in[r * b + c], s[r], p);
}

/* cpu_apply_rowmix_forward / _transpose, cuda/cuprym_cyclic.cu:1509-1566 */
static void rowmix(const orc_op* op, uint32_t* block, size_t b, int transpose) {
    uint32_t p = op->p;
    size_t pc = op->rows / 2;
    for (size_t t = 0; t < op->rounds; ++t) {
        size_t round = transpose ? op->rounds - 1 - t : t;
        const uint32_t* pr = op->pairs + round * pc * 4;
        for (size_t q = 0; q < pc; ++q) {
            size_t f = pr[4 * q], s = pr[4 * q + 1];
            uint64_t a = pr[4 * q + 2], bb = pr[4 * q + 3];
            /* (1 + ab) reduced first: for p < 2^31 all terms stay below 2^63 */
            uint64_t opab = (1 + (a * bb) % p) % p;
            for (size_t c = 0; c < b; ++c) {
                uint64_t x = block[f * b + c], y = block[s * b + c];
                uint64_t first = transpose ? x + bb * y : x + a * y;
                uint64_t second = transpose ? a * x + opab * y : bb * x + opab * y;
                block[f * b + c] = (uint32_t)(first % p);
                block[s * b + c] = (uint32_t)(second % p);
            }
        }
    }
}

/* y[s,a] = sum_kappa B[a][kappa] z[s,kappa] - sum_beta R[a][beta] t[s,beta]
 * (cpu_elim_z_to_y + cpu_elim_subtract_Rt, cuda/cuprym_cyclic.cu:1704-1777) */
static void elim_y(const orc_op* op, const uint32_t* z, const uint32_t* t, uint32_t* y,
                   size_t b) {
    uint32_t p = op->p;
    const phi_op* d2 = &op->d2;
    const phi_op* dm = &op->dm;
#pragma omp parallel for schedule(static)
    for (size_t row = 0; row < d2->cols; ++row) {
        uint32_t s = d2->col_rank[row], a = d2->col_a[row];
        for (size_t c = 0; c < b; ++c) {
            u128 plus = 0, minus = 0;
            for (uint32_t idx = op->z_off[s]; idx < op->z_off[s + 1]; ++idx)
                plus += (uint64_t)op->kernel_basis[(size_t)a * op->kernel_dim + op->z_kappa[idx]] * z[(size_t)idx * b + c];
            for (uint32_t idx = dm->row_off[s]; idx < dm->row_off[s + 1]; ++idx)
                minus += (uint64_t)op->right_inverse[(size_t)a * op->n_v + dm->row_beta[idx]] * t[(size_t)idx * b + c];
            uint64_t pv = (uint64_t)(plus % p), mv = (uint64_t)(minus % p);
            y[row * b + c] = (uint32_t)((pv + p - mv) % p);
        }
    }
}

/* forward F*(in): input already Dc-scaled */
static void forward_base(orc_op* op, const uint32_t* scaled, uint32_t* out, size_t b) {
    uint32_t p = op->p;
    if (op->kind == ORC_SYNTHETIC) {
        phi_op* d = &op->d_op;
        phi_op* e = &op->e_op;
        uint32_t* mid = (uint32_t*)calloc(d->rows * b + 1, sizeof(uint32_t));
        phi_forward(d, scaled, mid, b, p, 0, d->rows);
        phi_transpose(e, mid, out, b, p);   /* E = transpose gather with table E */
        free(mid);
    } else {
        size_t x_cols = op->x_cols;
        uint32_t* t = (uint32_t*)malloc((op->dm.rows * b + 1) * sizeof(uint32_t));
        uint32_t* y = (uint32_t*)malloc((op->d2.cols * b + 1) * sizeof(uint32_t));
        phi_forward(&op->dm, scaled, t, b, p, 0, op->dm.rows);
        elim_y(op, scaled + x_cols * b, t, y, b);
        phi_forward(&op->d2, y, out, b, p, 0, op->d2.rows);
        free(t);
        free(y);
    }
}

/* transpose F^T*(in), no Dc */
static void backward_base(orc_op* op, const uint32_t* in, uint32_t* out, size_t b) {
    uint32_t p = op->p;
    if (op->kind == ORC_SYNTHETIC) {
        phi_op* d = &op->d_op;
        phi_op* e = &op->e_op;
        uint32_t* mid = (uint32_t*)malloc((e->rows * b + 1) * sizeof(uint32_t));
        phi_forward(e, in, mid, b, p, 0, e->rows);  /* E^T = forward gather, table E */
        phi_transpose(d, mid, out, b, p);           /* D^T = transpose gather, table T */
        free(mid);
    } else {
        /* cpu_eliminated_backward, cuda/cuprym_cyclic.cu:1881-1900 */
        const phi_op* d2 = &op->d2;
        const phi_op* dm = &op->dm;
        uint32_t* h = (uint32_t*)malloc((d2->cols * b + 1) * sizeof(uint32_t));
        uint32_t* tadj = (uint32_t*)malloc((dm->rows * b + 1) * sizeof(uint32_t));
        phi_transpose(d2, in, h, b, p);
        /* t_adj[s,beta] = sum_a R[a][beta] h[s,a]  (cpu_elim_RT_h :1779-1806) */
        for (size_t row = 0; row < dm->rows; ++row) {
            uint32_t s = dm->row_rank[row], beta = dm->row_beta[row];
            for (size_t c = 0; c < b; ++c) {
                u128 sum = 0;
                for (uint32_t idx = d2->col_off[s]; idx < d2->col_off[s + 1]; ++idx)
                    sum += (uint64_t)op->right_inverse[(size_t)d2->col_a[idx] * op->n_v + beta] * h[(size_t)idx * b + c];
                tadj[row * b + c] = (uint32_t)(sum % p);
            }
        }
        phi_transpose(dm, tadj, out, b, p);  /* x_adj */
        /* z_adj[s,kappa] = sum_a B[a][kappa] h[s,a]  (cpu_elim_BT_h :1808-1835) */
        uint32_t* zadj = out + op->x_cols * b;
        for (size_t row = 0; row < op->z_count; ++row) {
            uint32_t s = op->z_rank[row], kappa = op->z_kappa[row];
            for (size_t c = 0; c < b; ++c) {
                u128 sum = 0;
                for (uint32_t idx = d2->col_off[s]; idx < d2->col_off[s + 1]; ++idx)
                    sum += (uint64_t)op->kernel_basis[(size_t)d2->col_a[idx] * op->kernel_dim + kappa] * h[(size_t)idx * b + c];
                zadj[row * b + c] = (uint32_t)(sum % p);
            }
        }
        /* out = [-x_adj ; z_adj]  (cpu_elim_assemble_adjoint :1837-1864) */
        for (size_t i = 0; i < op->x_cols * b; ++i) out[i] = out[i] ? p - out[i] : 0;
        free(h);
        free(tadj);
    }
}

/* base column of kept column j (replacement: the base columns in order with
 * drop0 and drop1 left out) */
static size_t kept_to_base(const orc_op* op, size_t j) {
    size_t f = j;
    if (f >= op->drop0) ++f;
    if (f >= op->drop1) ++f;
    return f;
}

/* G*(in) for the deformation variants (input already scaled) */
static void forward_core(orc_op* op, const uint32_t* scaled, uint32_t* out, size_t b) {
    if (!op->deform) {
        forward_base(op, scaled, out, b);
        return;
    }
    const uint32_t p = op->p;
    const size_t base = op->base_cols, dc = op->dense_cols;
    const size_t lead = op->deform == 1 ? base : base - 2;  /* columns before the dense block */
    const uint32_t* extra = scaled + lead * b;
    if (op->deform == 1) {
        forward_base(op, scaled, out, b);
    } else {
        uint32_t* full = (uint32_t*)calloc(base * b + 1, sizeof(uint32_t));
        for (size_t j = 0; j < lead; ++j)
            memcpy(full + kept_to_base(op, j) * b, scaled + j * b, b * sizeof(uint32_t));
        forward_base(op, full, out, b);
        free(full);
    }
    for (size_t r = 0; r < op->rows; ++r)
        for (size_t c = 0; c < b; ++c) {
            u128 sum = 0;
            for (size_t j = 0; j < dc; ++j) sum += (uint64_t)op->dense[r * dc + j] * extra[j * b + c];
            uint64_t t = (uint64_t)(sum % p);
            uint64_t v = op->deform == 1 ? out[r * b + c] + (p - t) : out[r * b + c] + t;
            out[r * b + c] = (uint32_t)(v % p);
        }
}

/* G^T*(in) for the deformation variants, no Dc */
static void backward_core(orc_op* op, const uint32_t* in, uint32_t* out, size_t b) {
    if (!op->deform) {
        backward_base(op, in, out, b);
        return;
    }
    const uint32_t p = op->p;
    const size_t base = op->base_cols, dc = op->dense_cols;
    const size_t lead = op->deform == 1 ? base : base - 2;
    if (op->deform == 1) {
        backward_base(op, in, out, b);
    } else {
        uint32_t* full = (uint32_t*)malloc((base * b + 1) * sizeof(uint32_t));
        backward_base(op, in, full, b);
        for (size_t j = 0; j < lead; ++j)
            memcpy(out + j * b, full + kept_to_base(op, j) * b, b * sizeof(uint32_t));
        free(full);
    }
    for (size_t j = 0; j < dc; ++j)
        for (size_t c = 0; c < b; ++c) {
            u128 sum = 0;
            for (size_t r = 0; r < op->rows; ++r) sum += (uint64_t)op->dense[r * dc + j] * in[r * b + c];
            uint32_t t = (uint32_t)(sum % p);
            out[(lead + j) * b + c] = op->deform == 1 ? (t ? p - t : 0) : t;
        }
}

void orc_apply_forward(orc_op* op, const uint32_t* in, uint32_t* out, size_t b) {
    uint32_t* scaled = (uint32_t*)malloc((op->cols * b + 1) * sizeof(uint32_t));
    scale_rows(in, op->dc, scaled, op->cols, b, op->p);
    forward_core(op, scaled, out, b);
    free(scaled);
}

void orc_apply_backward(orc_op* op, const uint32_t* in, uint32_t* out, size_t b) {
    backward_core(op, in, out, b);
    scale_rows(out, op->dc, out, op->cols, b, op->p);
}

void orc_apply_square(orc_op* op, const uint32_t* in, uint32_t* out, size_t b) {
    uint32_t p = op->p;
    uint32_t* scaled = (uint32_t*)malloc((op->cols * b + 1) * sizeof(uint32_t));
    uint32_t* target = (uint32_t*)malloc((op->rows * b + 1) * sizeof(uint32_t));
    scale_rows(in, op->dc, scaled, op->cols, b, p);
    forward_core(op, scaled, target, b);
    /* cpu_apply_target_preconditioner, cuda/cuprym_cyclic.cu:1687-1702 */
    if (op->rounds > 0) rowmix(op, target, b, 0);
    scale_rows(target, op->dr, target, op->rows, b, p);
    if (op->rounds > 0) rowmix(op, target, b, 1);
    backward_core(op, target, out, b);
    scale_rows(out, op->dc, out, op->cols, b, p);
    free(scaled);
    free(target);
}

void orc_sweep_partial(orc_op* op, int which, const uint32_t* in, uint32_t* out, size_t b,
                       size_t row0, size_t row1) {
    if (op->kind != ORC_SYNTHETIC) return;
    switch (which) {
        case 0: phi_forward(&op->d_op, in, out, b, op->p, row0, row1); break;          /* D   */
        case 1: phi_transpose_range(&op->e_op, in, out, b, op->p, row0, row1); break;  /* E   */
        case 2: phi_forward(&op->e_op, in, out, b, op->p, row0, row1); break;          /* E^T */
        case 3: phi_transpose_range(&op->d_op, in, out, b, op->p, row0, row1); break;  /* D^T */
        default: break;
    }
}

void orc_elementwise_sample(orc_op* op, uint32_t* block, size_t b, size_t rows) {
    const uint32_t p = op->p;
    const size_t r = rows < op->cols ? rows : op->cols;
    const size_t pc = op->rows / 2, qmax = r / 2 < pc ? r / 2 : pc;
    scale_rows(block, op->dc, block, r, b, p);                      /* Dc on the input */
    for (int transpose = 0; transpose < 2; ++transpose) {           /* P, then P^T */
        for (size_t t = 0; t < op->rounds; ++t) {
            const size_t round = transpose ? op->rounds - 1 - t : t;
            const uint32_t* pr = op->pairs + round * pc * 4;
            for (size_t q = 0; q < qmax; ++q) {
                const size_t f = pr[4 * q], s2 = pr[4 * q + 1];
                const uint64_t a = pr[4 * q + 2], bb = pr[4 * q + 3];
                const uint64_t opab = (1 + (a * bb) % p) % p;
                for (size_t c = 0; c < b; ++c) {
                    const uint64_t x = block[f * b + c], y = block[s2 * b + c];
                    block[f * b + c] = (uint32_t)((transpose ? x + bb * y : x + a * y) % p);
                    block[s2 * b + c] = (uint32_t)((transpose ? a * x + opab * y : bb * x + opab * y) % p);
                }
            }
        }
        if (!transpose) scale_rows(block, op->dr, block, r < op->rows ? r : op->rows, b, p);  /* Dr */
    }
    scale_rows(block, op->dc, block, r, b, p);                      /* Dc on the output */
}

void orc_set_threads(int threads) {
#ifdef _OPENMP
    omp_set_num_threads(threads > 0 ? threads : 1);
#else
    (void)threads;
#endif
}

size_t orc_sweep_rows(const orc_op* op, int which) {
    return (which == 0 || which == 2) ? op->d_op.rows : op->e_op.cols;
}

/* packed upper triangle: ut(i,j) = i(2b-i+1)/2 + (j-i), cuda/prym_cuda_helpers.h:49-55 */
void orc_gram_upper(const uint32_t* x, const uint32_t* y, size_t rows, size_t b,
                    uint32_t p, uint32_t* out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t i = 0; i < b; ++i)
        for (size_t j = i; j < b; ++j) {
            u128 sum = 0;
            for (size_t r = 0; r < rows; ++r) sum += (uint64_t)x[r * b + i] * y[r * b + j];
            out[i * (2 * b - i + 1) / 2 + (j - i)] = (uint32_t)(sum % p);
        }
}

size_t orc_sequence(orc_op* op, const uint32_t* v, size_t b, size_t max_len,
                    uint32_t* seq, uint32_t* curv) {
    size_t N = op->cols, ut = b * (b + 1) / 2;
    uint32_t* B = (uint32_t*)malloc((N * b + 1) * sizeof(uint32_t));
    uint32_t* D = (uint32_t*)malloc((N * b + 1) * sizeof(uint32_t));
    memcpy(B, v, N * b * sizeof(uint32_t));
    size_t pos = 0;
    orc_gram_upper(B, B, N, b, op->p, seq + ut * pos++);
    size_t rounds = (max_len - pos + 3) / 4;
    for (size_t round = 0; round < rounds; ++round) {
        orc_apply_square(op, B, D, b);
        orc_gram_upper(B, D, N, b, op->p, seq + ut * pos++);
        orc_gram_upper(D, D, N, b, op->p, seq + ut * pos++);
        orc_apply_square(op, D, B, b);
        orc_gram_upper(B, D, N, b, op->p, seq + ut * pos++);
        orc_gram_upper(B, B, N, b, op->p, seq + ut * pos++);
    }
    memcpy(curv, B, N * b * sizeof(uint32_t));
    free(B);
    free(D);
    return pos;
}

void orc_accumulate_generator(const uint32_t* B, const uint32_t* G, uint32_t* D,
                              size_t rows, size_t b, uint32_t p) {
    for (size_t r = 0; r < rows; ++r)
        for (size_t c = 0; c < b; ++c) {
            u128 sum = D[r * b + c];
            for (size_t i = 0; i < b; ++i) sum += (uint64_t)B[r * b + i] * G[c * b + i];
            D[r * b + c] = (uint32_t)(sum % p);
        }
}

void orc_kernel_vectors(orc_op* op, const uint32_t* v, size_t b, const uint32_t* gens,
                        size_t ngen, uint32_t* d_out, uint32_t* product_out) {
    size_t N = op->cols;
    uint32_t* B = (uint32_t*)malloc((N * b + 1) * sizeof(uint32_t));
    uint32_t* nxt = (uint32_t*)malloc((N * b + 1) * sizeof(uint32_t));
    memcpy(B, v, N * b * sizeof(uint32_t));
    memset(d_out, 0, N * b * sizeof(uint32_t));
    for (size_t k = 0; k < ngen; ++k) {
        orc_accumulate_generator(B, gens + k * b * b, d_out, N, b, op->p);
        if (k + 1 < ngen) {
            orc_apply_square(op, B, nxt, b);
            uint32_t* t = B; B = nxt; nxt = t;
        }
    }
    orc_apply_forward(op, d_out, product_out, b);
    free(B);
    free(nxt);
}

static uint32_t powmod(uint32_t base, uint32_t e, uint32_t p) {
    uint64_t r = 1, x = base % p;
    while (e) {
        if (e & 1) r = r * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

size_t orc_rank_mod_p(const uint32_t* m, size_t rows, size_t cols, uint32_t p,
                      uint32_t* pivot_cols) {
    uint32_t* a = (uint32_t*)malloc((rows * cols + 1) * sizeof(uint32_t));
    for (size_t i = 0; i < rows * cols; ++i) a[i] = m[i] % p;
    size_t rank = 0;
    for (size_t col = 0; col < cols && rank < rows; ++col) {
        size_t piv = rows;
        for (size_t r = rank; r < rows; ++r)
            if (a[r * cols + col]) { piv = r; break; }
        if (piv == rows) continue;
        if (piv != rank)
            for (size_t c = 0; c < cols; ++c) {
                uint32_t t = a[rank * cols + c];
                a[rank * cols + c] = a[piv * cols + c];
                a[piv * cols + c] = t;
            }
        uint32_t inv = powmod(a[rank * cols + col], p - 2, p);
        for (size_t c = col; c < cols; ++c) a[rank * cols + c] = mulmod(a[rank * cols + c], inv, p);
        for (size_t r = 0; r < rows; ++r) {
            if (r == rank || !a[r * cols + col]) continue;
            uint32_t f = a[r * cols + col];
            for (size_t c = col; c < cols; ++c) {
                uint32_t sub = mulmod(f, a[rank * cols + c], p);
                a[r * cols + c] = a[r * cols + c] >= sub ? a[r * cols + c] - sub : a[r * cols + c] + p - sub;
            }
        }
        if (pivot_cols) pivot_cols[rank] = (uint32_t)col;
        ++rank;
    }
    free(a);
    return rank;
}

static void write_line(FILE* f, const uint32_t* v, size_t count, size_t stride, uint32_t p) {
    for (size_t i = 0; i < count; ++i) {
        if (i) fputc(' ', f);
        fprintf(f, "%u", v[i * stride] % p);
    }
    fputc('\n', f);
}

/* save_all_data -> save_wdm_file_sym (cuda/cuprym_cyclic.cu:965-1007,
 * cuda/wdmfiles.h:164-222): the preconditioners are written as stored, the
 * vectors and the sequence as canonical residues, one column per line. */
int orc_write_wdm(const char* path, size_t rows, size_t cols, uint32_t p,
                  const uint32_t* dr, const uint32_t* dc, const uint32_t* v,
                  const uint32_t* curv, size_t b, const uint32_t* seq, size_t L) {
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    fprintf(f, "%zu %zu %u %zu %zu\n", rows, cols, p, L, b);
    for (size_t i = 0; i < rows; ++i) { if (i) fputc(' ', f); fprintf(f, "%u", dr[i]); }
    fputc('\n', f);
    for (size_t i = 0; i < cols; ++i) { if (i) fputc(' ', f); fprintf(f, "%u", dc[i]); }
    fputc('\n', f);
    for (size_t c = 0; c < b; ++c) write_line(f, v + c, cols, b, p);
    for (size_t c = 0; c < b; ++c) write_line(f, curv + c, cols, b, p);
    size_t ut = b * (b + 1) / 2;
    for (size_t e = 0; e < ut; ++e) write_line(f, seq + e, L, ut, p);
    fclose(f);
    return 0;
}

/* ---- BCW rank consumer, iterative restatement ----------------------------
 * third_party/bcw/src/sigma_basis.rs:75-110 (process_input_sequence),
 * 233-309 (basis), 312-324 (analyze_delta), 112-150 (generator list) and
 * modular_linalg.rs:71-125 (lsp_decomposition), with the order basis built one
 * coefficient at a time: at step k the residual coefficient
 * Delta = [x^k] (M G) is recomputed from scratch, then
 * M <- (D1 + x Dx) L^-1 pi M exactly as `basis` forms it. */

static void lsp_literal(uint32_t* a, uint32_t* l, size_t m, uint32_t p) {
    /* a: m x m in/out (becomes S with its columns permuted), l: m x m out */
    size_t* perm = (size_t*)malloc(m * sizeof(size_t));
    for (size_t i = 0; i < m; ++i) perm[i] = i;
    memset(l, 0, m * m * sizeof(uint32_t));
    for (size_t i = 0; i < m; ++i) l[i * m + i] = 1;
    size_t rank = 0;
    for (size_t k = 0; k < m; ++k) {
        size_t j = m;
        for (size_t c = rank; c < m; ++c)
            if (a[k * m + c] % p) { j = c; break; }
        if (j == m) continue;
        for (size_t r = 0; r < m; ++r) {
            uint32_t t = a[r * m + rank];
            a[r * m + rank] = a[r * m + j];
            a[r * m + j] = t;
        }
        size_t tp = perm[rank]; perm[rank] = perm[j]; perm[j] = tp;
        uint32_t inv = powmod(a[k * m + rank], p - 2, p);
        for (size_t i = k + 1; i < m; ++i) {
            uint32_t factor = mulmod(a[i * m + rank], inv, p);
            l[i * m + k] = factor;
            for (size_t col = rank; col < m; ++col) {
                uint32_t sub = mulmod(factor, a[k * m + col], p);
                a[i * m + col] = (a[i * m + col] + p - sub) % p;
            }
        }
        ++rank;
    }
    free(perm);
}

/* inverse of a unit lower-triangular matrix by forward substitution */
static void unit_lower_inverse(const uint32_t* l, uint32_t* inv, size_t m, uint32_t p) {
    memset(inv, 0, m * m * sizeof(uint32_t));
    for (size_t c = 0; c < m; ++c) {
        inv[c * m + c] = 1;
        for (size_t r = c + 1; r < m; ++r) {
            uint64_t s = 0;
            for (size_t t = c; t < r; ++t) s += (uint64_t)l[r * m + t] * inv[t * m + c] % p;
            inv[r * m + c] = (uint32_t)((p - s % p) % p);
        }
    }
}

int orc_bcw_rank(uint32_t p, size_t b, const uint32_t* seq, size_t L, uint64_t* delta_out,
                 uint64_t* rank_out, uint64_t* spread_out, uint32_t* gen, size_t gen_cap,
                 size_t* gen_terms) {
    if (L < 2 || b == 0) return -1;
    const size_t d = L - 1, m = 2 * b, n = b, ut = b * (b + 1) / 2, cap = d + 2;
    /* G[i][j][t], i < 2b, j < b */
    uint32_t* G = (uint32_t*)calloc(m * n * d, sizeof(uint32_t));
    size_t ii = 0;
    for (size_t i = 0; i < b; ++i)
        for (size_t j = i; j < b; ++j, ++ii)
            for (size_t t = 0; t < d; ++t) {
                G[(i * n + j) * d + t] = seq[(t + 1) * ut + ii] % p;
                G[(j * n + i) * d + t] = seq[(t + 1) * ut + ii] % p;
            }
    for (size_t i = 0; i < b; ++i) G[((b + i) * n + i) * d + 0] = 1;
    /* M[i][j][e], e < cap */
    uint32_t* M = (uint32_t*)calloc(m * m * cap, sizeof(uint32_t));
    uint32_t* T = (uint32_t*)calloc(m * m * cap, sizeof(uint32_t));
    for (size_t i = 0; i < m; ++i) M[(i * m + i) * cap] = 1;
    uint64_t* delta = (uint64_t*)malloc(m * sizeof(uint64_t));
    uint64_t* sdelta = (uint64_t*)malloc(m * sizeof(uint64_t));
    for (size_t i = 0; i < m; ++i) delta[i] = i < b ? 0 : 1;
    size_t* perm = (size_t*)malloc(m * sizeof(size_t));
    uint32_t* dm = (uint32_t*)malloc(m * n * sizeof(uint32_t));
    uint32_t* aug = (uint32_t*)malloc(m * m * sizeof(uint32_t));
    uint32_t* l = (uint32_t*)malloc(m * m * sizeof(uint32_t));
    uint32_t* linv = (uint32_t*)malloc(m * m * sizeof(uint32_t));
    char* shift = (char*)malloc(m);
    for (size_t k = 0; k < d; ++k) {
        /* Delta = [x^k] (M G) */
        for (size_t i = 0; i < m; ++i)
            for (size_t j = 0; j < n; ++j) {
                uint64_t s = 0;
                for (size_t t = 0; t < m; ++t)
                    for (size_t e = 0; e <= k && e < cap; ++e) {
                        uint32_t x = M[(i * m + t) * cap + e];
                        if (x) s += (uint64_t)x * G[(t * n + j) * d + (k - e)] % p;
                    }
                dm[i * n + j] = (uint32_t)(s % p);
            }
        /* stable sort of the row indices by delta (insertion sort) */
        for (size_t i = 0; i < m; ++i) perm[i] = i;
        for (size_t i = 1; i < m; ++i) {
            size_t x = perm[i], j = i;
            while (j > 0 && delta[perm[j - 1]] > delta[x]) { perm[j] = perm[j - 1]; --j; }
            perm[j] = x;
        }
        for (size_t i = 0; i < m; ++i) sdelta[i] = delta[perm[i]];
        /* aug = pi Delta padded to m x m; L S P = aug */
        memset(aug, 0, m * m * sizeof(uint32_t));
        for (size_t i = 0; i < m; ++i)
            for (size_t j = 0; j < n; ++j) aug[i * m + j] = dm[perm[i] * n + j];
        lsp_literal(aug, l, m, p);
        unit_lower_inverse(l, linv, m, p);
        for (size_t i = 0; i < m; ++i) {
            shift[i] = 0;
            for (size_t j = 0; j < m; ++j)
                if (aug[i * m + j]) { shift[i] = 1; break; }
        }
        /* T = L^-1 pi M, then rows with a nonzero S row times x */
        memset(T, 0, m * m * cap * sizeof(uint32_t));
        for (size_t i = 0; i < m; ++i)
            for (size_t t = 0; t < m; ++t) {
                uint32_t f = linv[i * m + t];
                if (!f) continue;
                const uint32_t* src = M + perm[t] * m * cap;
                uint32_t* dst = T + i * m * cap;
                for (size_t q = 0; q < m * cap; ++q)
                    if (src[q]) dst[q] = (uint32_t)((dst[q] + (uint64_t)f * src[q]) % p);
            }
        for (size_t i = 0; i < m; ++i) {
            for (size_t j = 0; j < m; ++j) {
                uint32_t* dst = M + (i * m + j) * cap;
                const uint32_t* src = T + (i * m + j) * cap;
                if (shift[i]) {
                    dst[0] = 0;
                    memcpy(dst + 1, src, (cap - 1) * sizeof(uint32_t));
                } else {
                    memcpy(dst, src, cap * sizeof(uint32_t));
                }
            }
            delta[i] = sdelta[i] + (shift[i] ? 1 : 0);
        }
    }
    /* analyze_delta */
    for (size_t i = 0; i < m; ++i) sdelta[i] = delta[i];
    for (size_t i = 1; i < m; ++i) {
        uint64_t x = sdelta[i];
        size_t j = i;
        while (j > 0 && sdelta[j - 1] > x) { sdelta[j] = sdelta[j - 1]; --j; }
        sdelta[j] = x;
    }
    uint64_t rank = 0;
    for (size_t i = 0; i < b; ++i) rank += sdelta[i];
    *rank_out = rank;
    *spread_out = sdelta[b] - sdelta[b - 1];
    for (size_t i = 0; i < m; ++i) delta_out[i] = delta[i];
    /* generator: row i < b reversed with its delta, lines after the last
     * nonzero dropped; line k row-major over (i, j) */
    size_t maxdeg = 0;
    for (size_t i = 0; i < b; ++i)
        for (size_t j = 0; j < b; ++j)
            for (size_t k = 0; k <= delta[i]; ++k)
                if (delta[i] - k < cap && M[(i * m + j) * cap + (delta[i] - k)] && k > maxdeg) maxdeg = k;
    *gen_terms = maxdeg + 1;
    int rc = 0;
    if (gen) {
        if (gen_cap < (maxdeg + 1) * b * b) rc = -2;
        else
            for (size_t k = 0; k <= maxdeg; ++k)
                for (size_t i = 0; i < b; ++i)
                    for (size_t j = 0; j < b; ++j)
                        gen[(k * b + i) * b + j] = (k <= delta[i] && delta[i] - k < cap)
                                                       ? M[(i * m + j) * cap + (delta[i] - k)] : 0;
    }
    free(G); free(M); free(T); free(delta); free(sdelta); free(perm); free(dm); free(aug);
    free(l); free(linv); free(shift);
    return rc;
}

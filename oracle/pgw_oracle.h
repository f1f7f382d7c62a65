/*
 * pgw_oracle.h -- CPU restatement of the reference's block Wiedemann sequence
 * path, used ONLY as a checker by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in the product (the HIP library, its
 * C ABI, the CLI) links, loads or calls this code.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository wilthoma/prym_green_ancillary).  Arithmetic is
 * exact: u64/u128 sums and one reduction per output, generalised from the
 * reference's uint16 storage to residues mod any prime p < 2^31.
 *
 * Parity pins (what checks each function, tests/test_oracle.py):
 *   - subsets/incidences: brute force over itertools.combinations order and
 *     the (-1)^position closed form;
 *   - sweeps / eliminated operator: dense brute-force assembly on tiny inputs,
 *     transpose identities, and the genus-12 certificate K0 (F0*K0 == 0),
 *     F^(0) nnz == 8400 and the sector ranks of g12-reference.json;
 *   - splitmix64 / FNV-1a / mt19937_64: published test vectors;
 *   - uniform_int (libstdc++ 11 Lemire draw): golden vectors written by g++;
 *   - Gram / sequence: closed form V^T S^t V against dense powers;
 *   - WDM text: the hand-made diag(1,2) file of the BCW CLI test;
 *   - seeded sector value, realized row-mix seed, row-mix pair lists, plan
 *     hash, the Dr/Dc/V draws and the deterministic blocks: golden vectors
 *     written by the reference's own host C++ (cuda/cuprym_cyclic.cu and
 *     cuda/cyclic_cuda_helpers.h line ranges compiled by oracle/ref_build.sh;
 *     tests/golden/ref_plan.json, ref_state.json, tests/test_reference_golden.py);
 *   - eliminated cyclic forward F (Dc x) and P^T Dr P F (Dc x): the reference's
 *     own CPU forward on every g12 sector and three g20 sectors
 *     (tests/golden/ref_forward_g12.json, ref_forward_g20.json);
 *   - orc_sweep_partial (the CPU baseline): its row ranges equal the dense
 *     D, E, E^T, D^T sweeps;
 *   - sector ranks: data/cases.json expected_base_ranks for every g20 and g22
 *     sector and g24 sectors 0 and 1 (through the product CLI, GPU tests);
 *   - deformation variants: the g12 certificate (tests/test_deformation.py):
 *     [S_c; I] in the kernel of [F0_c | -F0_c S_c] with rank 86 - 2 = 84
 *     (augmented_shape / augmented_nullity), rank 84 of [F0_0 without J | Z0]
 *     (completion_rank), transpose identities.
 * The reference's CPU transpose half (F^T) calls a __host__ __device__ helper
 * that needs the CUDA headers and is not built; F^T is pinned by the transpose
 * identities and by the product's independent host apply.  No reference WDM
 * file exists to compare bytes with; the WDM text is pinned by the BCW test
 * file above.
 */
#ifndef PGW_ORACLE_H
#define PGW_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- combinatorics (cuda/cuprym_cyclic.cu:134-230) ---- */
uint64_t orc_binom(uint32_t n, uint32_t k);
int orc_unrank_subset(uint64_t rank, uint32_t n, uint32_t k, uint32_t* out);

/* ---- RNG and hashes (cuda/cuprym_cyclic.cu:1143-1218, cuda/matrices.h:44-56) ---- */
uint64_t orc_splitmix64_next(uint64_t* state);
uint64_t orc_seeded_sector_value(uint64_t seed, uint64_t sector);
uint64_t orc_rowmix_bounded(uint64_t* state, uint64_t bound);
uint64_t orc_fnv1a64_u32_vector(const uint32_t* values, size_t count);
uint64_t orc_fnv1a64_string(const char* s);
/* mt19937_64 seeded with `seed`, first `count` raw outputs. */
void orc_mt19937_64_outputs(uint64_t seed, size_t count, uint64_t* out);
/* Draw order of run_eliminated_sector (cuda/cuprym_cyclic.cu:2196-2199):
 * Dr (nonzero, rows), Dc (nonzero, cols), V (cols*b, row-major). */
void orc_reference_random_state(uint64_t seed, uint64_t sector, uint32_t p,
                                size_t rows, size_t cols, size_t b,
                                uint32_t* dr, uint32_t* dc, uint32_t* v);
/* Uniform draws in [lo, hi] from mt19937_64(seed) (libstdc++ 11 semantics). */
void orc_uniform_u64(uint64_t seed, uint64_t lo, uint64_t hi, size_t count, uint64_t* out);

/* ---- row mixing (cuda/cuprym_cyclic.cu:1196-1287) ---- */
uint64_t orc_rowmix_realized_seed(uint64_t seed, const char* op, uint64_t sector,
                                  uint64_t rows, uint32_t p);
/* pairs: rounds * (rows/2) records of {first, second, a, b}. */
void orc_rowmix_plan(uint64_t realized_seed, size_t rows, size_t rounds, uint32_t p,
                     uint32_t* pairs);
uint64_t orc_rowmix_plan_hash(size_t rows, size_t rounds, uint64_t seed,
                              uint64_t realized_seed, const uint32_t* pairs);

/* ---- synthetic tables (uniform in [0,p) from a splitmix64 stream) ---- */
void orc_synthetic_table(uint64_t seed, uint32_t p, size_t count, uint32_t* out);

/* ---- operators ---- */
typedef struct orc_op orc_op;

/* Synthetic Koszul pair: D uses table T[i][beta][a] (forward gather),
 * E uses table E[i][beta][a] (transpose gather); F = E o D, square N x N,
 * N = C(n,k)*M, mid R = C(n,k-1)*Mp. */
orc_op* orc_synthetic_create(uint32_t n, uint32_t k, uint32_t M, uint32_t Mp, uint32_t p,
                             const uint32_t* T, const uint32_t* E,
                             const uint32_t* dr, const uint32_t* dc,
                             size_t rowmix_rounds, const uint32_t* rowmix_pairs);

/* Reference eliminated cyclic operator (cuda/cuprym_cyclic.cu:2136-2418). */
orc_op* orc_cyclic_create(uint32_t genus, uint32_t r, uint32_t p, uint32_t sector,
                          const uint32_t* mu, const uint16_t* weight_v,
                          const uint16_t* weight_a0, const uint16_t* weight_a1,
                          uint32_t w_index, uint32_t w_weight, uint32_t kernel_dim,
                          const uint16_t* kernel_weights, const uint32_t* right_inverse,
                          const uint32_t* kernel_basis);
/* Set preconditioners / rowmix after the dims are known (cyclic op). */
/* Deformation variants of an eliminated cyclic operator
 * (cuda/cuprym_deformation.cu:2554-2855): variant 1 augmented [F0 | -dense],
 * variant 2 replacement [F0 without base columns drop0 < drop1 | dense];
 * dense is rows x dense_cols (kept by reference).  Call before
 * orc_set_preconditioners; the operator's column count becomes that of G.
 * Returns 0, or -1 for an invalid request. */
int orc_cyclic_deformation(orc_op* op, int variant, const uint32_t* dense, size_t dense_cols,
                           size_t drop0, size_t drop1);
void orc_set_preconditioners(orc_op* op, const uint32_t* dr, const uint32_t* dc,
                             size_t rowmix_rounds, const uint32_t* rowmix_pairs);
void orc_op_dims(const orc_op* op, size_t* rows, size_t* cols);
void orc_op_destroy(orc_op* op);

/* F*(Dc*X) without row preconditioning (certificate product, cuprym_cyclic.cu:2330-2359). */
void orc_apply_forward(orc_op* op, const uint32_t* in, uint32_t* out, size_t b);
/* Dc*F^T*X (cuprym_cyclic.cu:2361-2395). */
void orc_apply_backward(orc_op* op, const uint32_t* in, uint32_t* out, size_t b);
/* S*X = Dc F^T P^T Dr P F Dc X (cuprym_cyclic.cu:2408-2418, 1939-1966). */
void orc_apply_square(orc_op* op, const uint32_t* in, uint32_t* out, size_t b);
/* One Koszul sweep of a synthetic operator restricted to output rows
 * [row0,row1): which = 0 the D sweep (forward gather, table T, N -> R),
 * 1 the E sweep (transpose gather, table E, R -> N), 2 the E^T sweep
 * (forward gather, table E, N -> R), 3 the D^T sweep (transpose gather,
 * table T, R -> N).  Used to time a bounded CPU sample of the apply; `out`
 * must hold the whole sweep output. */
void orc_sweep_partial(orc_op* op, int which, const uint32_t* in, uint32_t* out, size_t b,
                       size_t row0, size_t row1);
/* Timing sample of the elementwise stages of one apply (cpu_baseline only):
 * Dc scaling of the first `rows` rows, the row-mixing rounds (then their
 * transposes) on the pairs in the first rows/2 slots of each round, Dr and Dc
 * scaling.  `block` holds the operator's full row count x b and is scratch. */
void orc_elementwise_sample(orc_op* op, uint32_t* block, size_t b, size_t rows);
/* OpenMP threads of the oracle's parallel loops. */
void orc_set_threads(int threads);
/* Output rows of a synthetic sweep (R for D and E^T, N for E and D^T). */
size_t orc_sweep_rows(const orc_op* op, int which);

/* upper-triangle Gram, packed seq[ut(i,j)] (cuda/prym_cuda_helpers.h:87-137,318-354) */
void orc_gram_upper(const uint32_t* x, const uint32_t* y, size_t rows, size_t b,
                    uint32_t p, uint32_t* out);
/* Full sequence loop (cuprym_cyclic.cu:2326-2450): seq is L x b(b+1)/2
 * with L = 1 + 4*ceil((max_len-1)/4); curv is S^(2*rounds) V (two applies per
 * round of four terms). Returns L. */
size_t orc_sequence(orc_op* op, const uint32_t* v, size_t b, size_t max_len,
                    uint32_t* seq, uint32_t* curv);
/* D[r][c] += sum_i B[r][i]*G[c*b+i] (cuprym_cyclic_kernel_vectors.cu:1766-1790). */
void orc_accumulate_generator(const uint32_t* B, const uint32_t* G, uint32_t* D,
                              size_t rows, size_t b, uint32_t p);
/* Kernel-vector recovery core (cuprym_cyclic_kernel_vectors.cu:2549-2628):
 * D = sum_k S^k V G_k, product = F*(Dc*D).  */
void orc_kernel_vectors(orc_op* op, const uint32_t* v, size_t b, const uint32_t* gens,
                        size_t ngen, uint32_t* d_out, uint32_t* product_out);
/* Rank of a dense rows x cols matrix mod p (cuprym_cyclic_kernel_vectors.cu:1490-1546). */
size_t orc_rank_mod_p(const uint32_t* m, size_t rows, size_t cols, uint32_t p,
                      uint32_t* pivot_cols);

/* BCW rank consumer (third_party/bcw/src/sigma_basis.rs:75-150,233-324,
 * modular_linalg.rs:71-125), one order-basis step per coefficient, residual
 * recomputed from M G each step.  seq: L x b(b+1)/2 (term 0 = V^T V, dropped).
 * Outputs the 2b shifted degrees (basis row order), the rank lower bound, the
 * gap, and the generator matrices (gen[k*b*b + i*b + j], *gen_terms lines).
 * Returns 0, -1 for L < 2, -2 when gen_cap is too small. */
int orc_bcw_rank(uint32_t p, size_t b, const uint32_t* seq, size_t L, uint64_t* delta_out,
                 uint64_t* rank_out, uint64_t* spread_out, uint32_t* gen, size_t gen_cap,
                 size_t* gen_terms);

/* WDM text writer (cuda/wdmfiles.h:153-222), uncompressed. seq is L x b(b+1)/2. */
int orc_write_wdm(const char* path, size_t rows, size_t cols, uint32_t p,
                  const uint32_t* dr, const uint32_t* dc, const uint32_t* v,
                  const uint32_t* curv, size_t b, const uint32_t* seq, size_t L);

#ifdef __cplusplus
}
#endif
#endif

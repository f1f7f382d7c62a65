/*
 * pgw.h -- C ABI of the MI355X block Wiedemann sequence generator (libpgw.so).
 *
 * The reference (wilthoma/prym_green_ancillary) has no FFI: its boundary is the
 * `cuprym_cyclic` / `cuprym_cyclic_kernel_vectors` executables plus the
 * header-only C++ operator API compiled into them.  Each entry point below
 * names the reference interface it replaces.  Conventions:
 *   - every call returns PGW_OK (0) or a nonzero pgw_status; no exception ever
 *     crosses the ABI; pgw_last_error() gives the message (thread-local);
 *   - blocks are row-major N x b arrays of uint32 canonical residues
 *     (X[row*b + col]), exactly the reference's dense layout
 *     (cuda/cyclic_cuda_helpers.h:185, X[(colbase+a)*DENSE_COLS + c]);
 *   - d_* pointers are device memory owned by the caller, h_* host memory;
 *   - `stream` is a hipStream_t (NULL = the context's own stream); calls on
 *     one operator are serialised on the stream they are given;
 *   - an operator is thread-compatible, not thread-safe.
 */
#ifndef PGW_H
#define PGW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pgw_ctx pgw_ctx;
typedef struct pgw_op pgw_op;

enum pgw_status {
    PGW_OK = 0,
    PGW_ERR_INVALID = 1,      /* bad argument or shape */
    PGW_ERR_HIP = 2,          /* HIP runtime / launch failure */
    PGW_ERR_MEMORY = 3,       /* device or host allocation failed */
    PGW_ERR_UNSUPPORTED = 4,  /* valid request outside what this build runs */
    PGW_ERR_IO = 5,           /* file I/O (WDM writer / reader) */
    PGW_ERR_NEED_MORE_SEQUENCE = 6  /* rank step: no terms beyond those processed */
};

enum pgw_kind {
    /* A = E o D over Lambda^k V (x) R_d, dense multiplication tables (r = 1). */
    PGW_KIND_SYNTHETIC = 1,
    /* the reference's eliminated cyclic-sector operator
     * F = [ -D_(m-1) R D_m | D_(m-1) B ]  (cuda/cuprym_cyclic.cu:2330-2359) */
    PGW_KIND_CYCLIC_ELIMINATED = 2,
    /* deformation variants of that operator (run_deformation_eliminated_variant,
     * cuda/cuprym_deformation.cu:2554-2855): augmented G = [F | -dense],
     * replacement G = [F without base columns drop_cols[0] < drop_cols[1] | dense];
     * source columns: base (or kept base) columns, then the dense columns. */
    PGW_KIND_CYCLIC_AUGMENTED = 3,
    PGW_KIND_CYCLIC_REPLACEMENT = 4
};

enum pgw_method {
    PGW_METHOD_AUTO = 0,   /* MFMA where the operator allows it, else plain */
    PGW_METHOD_PLAIN = 1,  /* u64 modmul gathers on the vector ALUs */
    PGW_METHOD_MFMA = 2    /* int8-limb MFMA table contraction (synthetic only) */
};

/* One 2x2 row-mixing map [[1,a],[b,1+ab]] on target rows (first, second).
 * Replaces CyclicRowMixPair (cuda/cyclic_cuda_helpers.h:39-44), widened to
 * u32 coefficients so that p may exceed 65535. */
typedef struct {
    uint32_t first;
    uint32_t second;
    uint32_t a;
    uint32_t b;
} pgw_rowmix_pair;

/* Operator description; arrays are deep-copied by pgw_operator_create.
 * Replaces CyclicPhiHostData + the elimination arrays of run_eliminated_sector
 * (cuda/cyclic_cuda_helpers.h:46-75, cuda/cuprym_cyclic.cu:2136-2298). */
typedef struct {
    uint32_t kind;        /* pgw_kind */
    uint32_t prime;       /* p, prime, 2 < p < 2^31 */
    uint32_t block;       /* b, block width, 1..16 */
    uint32_t method;      /* pgw_method */
    /* PGW_KIND_SYNTHETIC: D: Lambda^k (x) R_d -> Lambda^(k-1) (x) R_(d+1) with
     * table_d[i][beta][a]; E: Lambda^(k-1) (x) R_(d+1) -> Lambda^k (x) R_d with
     * table_e[i][beta][a]; i < n, beta < m_tgt, a < m_src. */
    uint32_t n;
    uint32_t k;
    uint32_t m_src;       /* M  = dim R_d   */
    uint32_t m_tgt;       /* M' = dim R_(d+1) */
    const uint32_t* table_d;
    const uint32_t* table_e;
    /* PGW_KIND_CYCLIC_ELIMINATED: the instance JSON fields
     * (cuda/cuprym_cyclic.cu:483-609); n = genus - 3. */
    uint32_t genus;
    uint32_t order;       /* r, cyclic order */
    uint32_t sector;
    const uint32_t* mu;   /* [n][n][genus], layout ((i*n)+beta)*genus+a */
    const uint16_t* weight_v;
    const uint16_t* weight_a0;
    const uint16_t* weight_a1;
    uint32_t w_index;
    uint32_t w_weight;
    uint32_t kernel_dim;
    const uint16_t* kernel_weights;
    const uint32_t* right_inverse;   /* [genus][n]          */
    const uint32_t* kernel_basis;    /* [genus][kernel_dim] */
    /* PGW_KIND_CYCLIC_AUGMENTED / _REPLACEMENT (the --dense-file and
     * --drop-columns inputs, cuda/cuprym_deformation.cu:2600-2619): the dense
     * block, target rows x dense_cols row-major, dense_cols = 2 as the reference
     * requires; the two base columns the replacement drops, ascending. */
    uint32_t dense_cols;
    const uint32_t* dense_block;
    uint32_t drop_cols[2];
} pgw_operator_desc;

typedef struct {
    uint64_t rows;          /* target rows of F */
    uint64_t cols;          /* source columns of F = N, the block height */
    uint64_t mid_rows;      /* intermediate space (R for synthetic, y for cyclic) */
    uint64_t device_bytes;  /* device memory the operator holds */
    uint32_t method;        /* resolved pgw_method */
    uint32_t block;
} pgw_op_info;

/* ---- errors and contexts ------------------------------------------------ */

/* Message of the last failed call on this thread (never NULL). */
const char* pgw_last_error(void);
/* Replaces cudaSetDevice + the legacy default stream (cuprym_cyclic.cu:2578). */
int pgw_ctx_create(int device, pgw_ctx** out);
int pgw_ctx_destroy(pgw_ctx* ctx);
int pgw_ctx_synchronize(pgw_ctx* ctx);
/* Device name and total memory (cuprym_cyclic.cu:2579-2583). */
int pgw_ctx_device_info(pgw_ctx* ctx, char* name, size_t name_len, uint64_t* total_bytes);

/* ---- host-side recipe helpers (same bits as the reference driver) -------- */

/* seeded_sector_value (cuda/cuprym_cyclic.cu:1143-1151). */
uint64_t pgw_seeded_sector_value(uint64_t seed, uint64_t sector);
/* mt19937_64(seeded_sector_value(seed, sector)) draws Dr (rows, nonzero),
 * Dc (cols, nonzero), V (cols*b) in that order (cuda/cuprym_cyclic.cu:2196-2199,
 * cuda/matrices.h:44-56). */
int pgw_reference_random_state(uint64_t seed, uint64_t sector, uint32_t prime, uint64_t rows,
                               uint64_t cols, uint32_t block, uint32_t* h_dr, uint32_t* h_dc,
                               uint32_t* h_v);
/* rowmix_realized_seed (cuda/cuprym_cyclic.cu:1196-1205). */
uint64_t pgw_rowmix_realized_seed(uint64_t seed, const char* op_name, uint64_t sector,
                                  uint64_t rows, uint32_t prime);
/* generate_rowmix_plan (cuda/cuprym_cyclic.cu:1227-1268): rounds*(rows/2) pairs. */
int pgw_rowmix_plan(uint64_t realized_seed, uint64_t rows, uint32_t rounds, uint32_t prime,
                    pgw_rowmix_pair* h_pairs);
/* rowmix_plan_hash (cuda/cuprym_cyclic.cu:1270-1287). */
uint64_t pgw_rowmix_plan_hash(uint64_t rows, uint32_t rounds, uint64_t seed,
                              uint64_t realized_seed, const pgw_rowmix_pair* h_pairs);
/* fnv1a64_vector_hash over u64 little-endian (cuda/cuprym_cyclic.cu:1181-1188). */
uint64_t pgw_fnv1a64_u32(const uint32_t* h_values, uint64_t count);
/* Synthetic multiplication table: `count` values uniform in [0,p) from a
 * splitmix64 stream with rejection-bounded draws (rowmix_bounded,
 * cuda/cuprym_cyclic.cu:1207-1218). No counterpart table exists in the
 * reference; the draw procedure is the reference's. */
int pgw_synthetic_table(uint64_t seed, uint32_t prime, uint64_t count, uint32_t* h_out);

/* ---- operators ------------------------------------------------------------ */

/* Shape of the operator a description defines, without touching the GPU
 * (build_sector_layout_generic / build_tensor_layout sizes,
 * cuda/cuprym_cyclic.cu:2161-2176). */
int pgw_operator_dims(const pgw_operator_desc* desc, uint64_t* rows, uint64_t* cols);
/* Replaces CudaCyclicPhiOperator::from_host x2 + the device copies of the
 * elimination data, preconditioners and row-mix plan (cuprym_cyclic.cu:2240-2324,
 * cuda/cyclic_cuda_helpers.h:297-360).  h_row_precond has `rows` entries (Dr),
 * h_col_precond `cols` entries (Dc), h_pairs rowmix_rounds*(rows/2) entries. */
int pgw_operator_create(pgw_ctx* ctx, const pgw_operator_desc* desc,
                        const uint32_t* h_row_precond, const uint32_t* h_col_precond,
                        uint32_t rowmix_rounds, const pgw_rowmix_pair* h_pairs, pgw_op** out);
/* Replaces CudaCyclicPhiOperator::release (cuda/cyclic_cuda_helpers.h:362-394). */
int pgw_operator_destroy(pgw_op* op);
int pgw_operator_info(const pgw_op* op, pgw_op_info* info);

/* One block-apply d_out = S*d_in, S = Dc F^T P^T Dr P F Dc, both N x b.
 * Replaces the apply_square lambda (cuda/cuprym_cyclic.cu:2408-2418). */
int pgw_apply_square(pgw_op* op, const uint32_t* d_in, uint32_t* d_out, void* stream);
/* d_out (rows x b) = F*(Dc*d_in): the certificate product
 * (apply_forward, cuda/cuprym_cyclic.cu:2330-2359; kernel_vectors.cu:2628). */
int pgw_apply_forward(pgw_op* op, const uint32_t* d_in, uint32_t* d_out, void* stream);
/* Packed upper triangle of d_x^T d_y mod p into d_out[b(b+1)/2], WDM order
 * ut(i,j) = i(2b-i+1)/2 + (j-i).  Replaces mTm_tri_pack_from_scratch
 * (cuda/prym_cuda_helpers.h:318-354) without its u32 atomic accumulators. */
int pgw_gram_upper(pgw_op* op, const uint32_t* d_x, const uint32_t* d_y, uint32_t* d_out,
                   void* stream);
/* Where the operator's block projections (pgw_gram_upper, the sequence loop,
 * ungrouped batch and server jobs) run: PGW_METHOD_MFMA on int8 MFMA (digits
 * of both blocks in registers, int32 digit-shift sums folded to residues
 * every 256 chunks of 64 rows; b in {4, 8, 16}), PGW_METHOD_PLAIN on the
 * vector ALUs (64-bit word sums), PGW_METHOD_AUTO the default: MFMA for
 * operators on the MFMA path.  The reference's projection is the u32 GEMM
 * dense_gemm_TN_chunked3D_offset_u32 (cuda/prym_cuda_helpers.h:107-137).
 * Results are identical for every choice. */
int pgw_operator_set_gram(pgw_op* op, uint32_t method);
/* Synthetic operators with one or two row-mixing rounds: build, once, the
 * chains the rounds link (host work linear in the rows: about 14 s at 53.8 M
 * rows) so that every later apply runs P^T Dr P in one pass
 * (k_mix_chain) instead of 2R - 1 passes over the pairs; the results are
 * the same.  A no-op for other operators, for 2^31 - 1 rows or more, and when
 * already prepared.  Call it while the operator has no work in flight;
 * pgw_sequence_run calls it itself for loops of 1024 rounds or more (and keeps
 * the pair passes if it fails).  No counterpart in the reference (its
 * row mixing is the pair passes, cuda/cyclic_cuda_helpers.h:109-143). */
int pgw_operator_prepare_row_mixing(pgw_op* op);

/* pgw_gram_upper without an operator and without any allocation: d_out
 * (b(b+1)/2 values) = packed upper triangle of d_x^T d_y mod `prime` for two
 * rows x block device blocks (1 <= block <= 16, rows < 2^32).  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its partial sums live in a fixed 2 MiB device array per GPU
 * (set up by pgw_ctx_create, held until the process ends); calls that
 * overlap on one GPU share that array and race, so serialise them. */
int pgw_gram_upper_noalloc(pgw_ctx* ctx, uint32_t prime, uint32_t block, uint64_t rows,
                           const uint32_t* d_x, const uint32_t* d_y, uint32_t* d_out,
                           void* stream);

/* The exact host apply of a cyclic operator (eliminated or a deformation
 * variant): h_out = S h_in, S = Dc F^T P^T Dr P F Dc, on a cols x width host
 * block, with the same description, preconditioners and row-mixing plan that
 * pgw_operator_create takes.  Replaces cpu_apply_eliminated_square
 * (cuda/cuprym_cyclic.cu:1591-1966): matrix-free from the sector layout in
 * the reference's step order, on the host only (no device is touched).  The
 * CLI's --validate-rowmix compares the device apply with it entry by entry
 * (validate_eliminated_cpu_cuda, cuda/cuprym_cyclic.cu:1970-2132). */
int pgw_cyclic_apply_cpu(const pgw_operator_desc* desc, const uint32_t* h_row_precond,
                         const uint32_t* h_col_precond, uint32_t rowmix_rounds,
                         const pgw_rowmix_pair* h_pairs, const uint32_t* h_in, uint32_t width,
                         uint32_t* h_out);
/* deterministic_dense_block (cuda/cuprym_cyclic.cu:1410-1420): rows x cols
 * residues mod prime, row-major, from the splitmix64 stream seeded with
 * salt ^ (0xD6E8FEB86659FD93 + rows * 1315423911 + cols * 2654435761). */
int pgw_deterministic_block(uint64_t rows, uint64_t cols, uint32_t prime, uint64_t salt,
                            uint32_t* h_out);

/* Progress callback: terms done, terms total. */
typedef void (*pgw_progress_fn)(uint64_t done, uint64_t total, void* user);

/* Whole sequence loop (cuprym_cyclic.cu:2316-2450): seq[0] = V^T V, then
 * ceil((max_len-1)/4) rounds of {D=S B; B^T D; D^T D; B=S D; B^T D; B^T B}.
 * h_seq receives L x b(b+1)/2 values with L = 1+4*ceil((max_len-1)/4)
 * (row t = packed V^T S^t V), h_curv the final N x b block.  *out_len = L.
 * Every 64 rounds the terms finished so far are copied to h_seq before
 * progress(done, L, user) is called, so rows [0, done) of h_seq hold their
 * final values inside the callback (a streaming WDM writer reads them while
 * the loop goes on; the host memory of later rows is written only after). */
int pgw_sequence_run(pgw_op* op, const uint32_t* h_v, uint64_t max_len, uint32_t* h_seq,
                     uint32_t* h_curv, uint64_t* out_len, pgw_progress_fn progress,
                     void* user);

/* Kernel-vector recovery core (cuprym_cyclic_kernel_vectors.cu:2549-2628):
 * h_d = sum_{k<ngen} S^k V G_k (G_k[c*b+i] multiplies column i into c),
 * h_product = F*(Dc*h_d) (rows x b, all zero for a certificate). */
int pgw_kernel_vectors(pgw_op* op, const uint32_t* h_v, const uint32_t* h_gens, uint64_t ngen,
                       uint32_t* h_d, uint32_t* h_product);

/* Autotuner: times every MFMA tile configuration of each of the four sweeps
 * (`reps` launches each, HIP events) and keeps the fastest for this operator.
 * h_choice[4] / h_ms[4] (optional) receive the choices and their per-launch
 * milliseconds.  No counterpart in the reference (its kernels have one
 * fixed launch shape, cuda/cyclic_cuda_helpers.h:455-489); results do not
 * depend on the choice. */
int pgw_operator_autotune(pgw_op* op, uint32_t reps, uint32_t* h_choice, double* h_ms);

/* Timing of the hot path on the operator's own stream with HIP events:
 * `iters` block-applies d_out = S d_in (ping-pong), total milliseconds and
 * the per-launch average of the dominant sweep kernel (ms). */
int pgw_bench_apply(pgw_op* op, uint32_t* d_a, uint32_t* d_b, uint32_t iters,
                    double* total_ms, double* sweep_avg_ms, uint32_t* sweeps_per_apply);

/* Stages of a block-apply for pgw_bench_stages. */
enum pgw_stage {
    PGW_STAGE_SWEEP = 0,      /* table-contraction sweeps (MFMA or plain) */
    PGW_STAGE_DIGITS = 1,     /* u32 block -> int8 digit planes (MFMA path) */
    PGW_STAGE_ROWMIX = 2,     /* P^T Dr P on the target block */
    PGW_STAGE_SCALE = 3,      /* Dc row scaling (plain path) */
    PGW_STAGE_SPARSE = 4      /* sparse maps of the cyclic operator */
};
#define PGW_STAGES 5
/* Per-stage timing of the apply on the operator's own stream: `iters`
 * block-applies (ping-pong between d_a and d_b), HIP events around every
 * launch.  h_ms[s] = mean milliseconds per apply in stage s, h_bytes[s] = the
 * stage's algorithmic HBM bytes per apply (unique input + output bytes),
 * h_launches[s] = its launches per apply; arrays of PGW_STAGES entries.  No
 * counterpart in the reference (it times only the whole loop,
 * cuda/cuprym_cyclic.cu:2420-2461). */
int pgw_bench_stages(pgw_op* op, uint32_t* d_a, uint32_t* d_b, uint32_t iters, double* h_ms,
                     double* h_bytes, uint32_t* h_launches);

/* ------------------------------------------------------------------------
 * Batches: many sectors' sequence loops advanced together.
 *
 * The reference runs one sector per process, one after another
 * (scripts/reproduce.py:235-276 launching cuda/cuprym_cyclic.cu:2316-2461
 * per sector).  A batch holds one job per operator (each job is exactly the
 * pgw_sequence_run loop of that operator); one replay advances every
 * unfinished job by one round (two block-applies, four projections), all on
 * the batch's single stream.  With PGW_BATCH_GROUPED, jobs of the synthetic
 * operator on the MFMA path that share block width, limb count, modulus
 * class, row-mixing rounds and tile configurations (7-9) form a group, and
 * every stage of the round is ONE launch for the whole group (job j owns a
 * block range of each grid); other jobs run their own launches.  With
 * PGW_BATCH_GRAPH a round is captured once into a single-stream HIP graph
 * and replayed while the set of jobs is unchanged.  Sequence positions live
 * in a device counter per job, so the same launches serve every round.  Each
 * job's sequence and final block are identical to pgw_sequence_run on the
 * same operator and V.  pgw_batch_add / _poll / _result may be called from
 * another thread than pgw_batch_step (jobs are staged on a stream of their
 * own).
 * ---------------------------------------------------------------------- */
typedef struct pgw_batch pgw_batch;

enum pgw_batch_flags {
    PGW_BATCH_GRAPH = 1,    /* replay a captured graph of the round (else launch every replay) */
    PGW_BATCH_GROUPED = 2,  /* one launch per stage for each group of compatible jobs */
    PGW_BATCH_AUTOTUNE = 4,  /* time tile configurations 7-9 once per operator shape (n, k, M, M',
                                b, limbs, fold) and give every job of that shape the fastest */
    PGW_BATCH_PERSISTENT = 8  /* grouped sweeps as persistent launches: two workgroups per CU walk
                                 every tile of the group (default: one workgroup per tile, which
                                 measured faster on the 64-sector mix) */
};

int pgw_batch_create(pgw_ctx* ctx, uint32_t flags, pgw_batch** out);
/* Adds a job: h_v (N x b) is copied, seq[0] = V^T V is computed at once.
 * An operator may carry at most one unfinished job.  *job receives its id. */
int pgw_batch_add(pgw_batch* b, pgw_op* op, const uint32_t* h_v, uint64_t max_len,
                  uint64_t* job);
/* Runs up to max_replays (0 = 1) replays, fewer where a job would pass its
 * end; *active (optional) receives the number of unfinished jobs.  Returns
 * when the replays are done. */
int pgw_batch_step(pgw_batch* b, uint32_t max_replays, uint64_t* active);
int pgw_batch_poll(pgw_batch* b, uint64_t job, int* done);
/* Copies a finished job's L x b(b+1)/2 sequence and final N x b block to
 * the host (same layout as pgw_sequence_run), frees its device state and
 * forgets the job. */
int pgw_batch_result(pgw_batch* b, uint64_t job, uint32_t* h_seq, uint32_t* h_curv,
                     uint64_t* len);
/* The terms of a job that are final while it still runs: *done = 1 + 4 *
 * (rounds of replays already synchronised), at most L, and with h_seq the
 * rows [first, *done) are copied to h_seq + first * b(b+1)/2 (h_seq sized
 * for all L rows; NULL only reports *done).  The copy runs on a stream of
 * its own, so a pgw_batch_step in progress is not waited for.  It must not
 * overlap pgw_batch_result, _cancel or _destroy of the same job.  No
 * counterpart in the reference; a serving loop feeds these terms to the rank
 * stage during the job. */
int pgw_batch_terms(pgw_batch* b, uint64_t job, uint64_t first, uint32_t* h_seq, uint64_t* done);
/* Drops a job (finished or not) and frees its device state. */
int pgw_batch_cancel(pgw_batch* b, uint64_t job);
/* Number of graph captures so far (instrumentation). */
int pgw_batch_graph_builds(pgw_batch* b, uint64_t* builds);
/* Groups of the current plan and the jobs they hold (instrumentation). */
int pgw_batch_groups(pgw_batch* b, uint64_t* groups, uint64_t* grouped_jobs);
int pgw_batch_destroy(pgw_batch* b);

/* ------------------------------------------------------------------------
 * Serving: sector jobs admitted by their device-memory footprint and run
 * as one continuously refilled batch.
 *
 * Replaces the reference's resource-pool scheduling of sector processes
 * (scripts/reproduce.py:78-111, 246-279) for sector jobs.  Admission is
 * FIFO: the head of the queue is admitted when its footprint fits in
 * `hbm_budget` bytes next to the running jobs and fewer than `max_active`
 * jobs run; a job whose footprint exceeds the whole budget fails at once.
 * Between batch steps (grouped, graph-replayed rounds) finished jobs leave
 * and admitted jobs join.  Results of a job do not depend on what else runs
 * with it.
 * ---------------------------------------------------------------------- */
typedef struct pgw_server pgw_server;

/* Device bytes a job is admitted with: operator (tables, digit planes,
 * intermediate blocks, preconditioners, row-mix plan, Gram scratch) plus its
 * sequence state (two N x b blocks, L x b(b+1)/2 terms).  An upper bound for
 * synthetic operators; once admitted, a job is charged what its operator
 * really holds. */
int pgw_job_footprint(const pgw_operator_desc* desc, uint64_t max_len, uint32_t rowmix_rounds,
                      uint64_t* bytes);
int pgw_server_create(pgw_ctx* ctx, uint64_t hbm_budget, uint32_t max_active,
                      uint32_t rounds_per_step, pgw_server** out);
/* Queues one sector job: the operator description and its arrays, Dr, Dc,
 * the row-mix plan and V are deep-copied (same meaning as
 * pgw_operator_create + pgw_sequence_run).  *ticket receives its id. */
int pgw_server_submit(pgw_server* s, const pgw_operator_desc* desc, const uint32_t* h_dr,
                      const uint32_t* h_dc, uint32_t rowmix_rounds,
                      const pgw_rowmix_pair* h_pairs, const uint32_t* h_v, uint64_t max_len,
                      uint64_t* ticket);
/* Waits up to timeout_s seconds (negative: no limit); *state receives
 * 0 queued, 1 running, 2 done, 3 failed. */
int pgw_server_wait(pgw_server* s, uint64_t ticket, double timeout_s, int* state);
/* Copies a finished job's sequence (L x b(b+1)/2) and final block (N x b)
 * (either pointer may be null) and forgets the ticket.  times (optional, 5
 * entries) receives, in seconds since the server started: submitted,
 * admitted, in the batch (operator built, V uploaded), out of the batch
 * (last round done), finished (results in host memory).  A failed job
 * returns PGW_ERR_INVALID with its error message. */
int pgw_server_result(pgw_server* s, uint64_t ticket, uint32_t* h_seq, uint32_t* h_curv,
                      uint64_t* len, double* times);
/* completed counts finished and failed jobs. */
int pgw_server_stats(pgw_server* s, uint64_t* completed, uint64_t* peak_bytes,
                     uint64_t* peak_active, uint64_t* queued);
/* Finishes every queued and running job, then stops the worker. */
int pgw_server_destroy(pgw_server* s);

/* ---- rank consumer (the reader of the WDM sequence) --------------------- */

/* Replaces the BCW stage `prym-rank` (third_party/bcw/src/rank_pipeline.rs:
 * 105-287, sigma_basis.rs:37-324): the rank lower bound of the Gram operator
 * from its symmetric sequence via a shifted order basis of [S(x); I_b] with
 * shift (0^b, 1^b), and the matrix generator (its first b rows, reversed).
 * Host code; the same basis as the reference's, for any prime p < 2^31. */
typedef struct pgw_rank pgw_rank;

/* RankComputationStatus (rank_pipeline.rs:42-47,236-246) */
enum pgw_rank_status {
    PGW_RANK_RUNNING = 0,             /* gap below the threshold, more terms may help */
    PGW_RANK_SUCCESS = 1,             /* gap >= threshold: the rank is final */
    PGW_RANK_EXCEEDS_DIMENSION = 2,   /* sum of the b smallest degrees > min(rows, cols) */
    PGW_RANK_SEQUENCE_LIMIT = 3       /* processed >= 2 min(rows, cols) / b + buffer */
};

/* RankComputationOptions::default() (rank_pipeline.rs:15-16,26-39) */
#define PGW_RANK_DEFAULT_THRESHOLD 10u
#define PGW_RANK_DEFAULT_MAX_NLEN_BUFFER 40u

typedef struct {
    uint64_t rank;          /* rank lower bound: sum of the b smallest shifted degrees */
    uint64_t delta_spread;  /* gap between the b-th and (b+1)-th smallest degree */
    uint64_t processed;     /* usable terms consumed (the initial term excluded) */
    uint64_t max_terms;     /* max_nlen = 2 min(rows, cols) / b + buffer */
    int status;             /* pgw_rank_status */
} pgw_rank_info;

/* RankComputation::new (rank_pipeline.rs:119-151).  rows x cols is the size of
 * F (the WDM header's m n); threads = 0 picks the host's cores (at most 16). */
int pgw_rank_create(uint32_t prime, uint32_t block, uint64_t n_rows, uint64_t n_cols,
                    uint32_t threshold, uint32_t max_nlen_buffer, uint32_t threads,
                    pgw_rank** out);
/* RankComputation::step (rank_pipeline.rs:153-222): h_seq holds `len` terms of
 * b(b+1)/2 packed upper-triangle values (term 0 = V^T V, dropped), the layout
 * of pgw_sequence_run and of the WDM file.  Consumes the terms after those
 * already processed; PGW_ERR_NEED_MORE_SEQUENCE when there are none. */
int pgw_rank_step(pgw_rank* r, const uint32_t* h_seq, uint64_t len, pgw_rank_info* info);
int pgw_rank_get_info(const pgw_rank* r, pgw_rank_info* info);
/* The 2b shifted degrees, in the basis' row order. */
int pgw_rank_delta(const pgw_rank* r, uint64_t* h_delta);
/* order_basis_to_generator_matrixlist + save_generator_list
 * (sigma_basis.rs:112-168): *terms b x b matrices, matrix k row-major at
 * h_out[k*b*b], exactly the lines of `<stem>_generators.txt`.  h_out = NULL
 * only reports *terms. */
int pgw_rank_generator(const pgw_rank* r, uint32_t* h_out, uint64_t capacity, uint64_t* terms);
int pgw_rank_destroy(pgw_rank* r);

/* Host rank stage of a server (no counterpart in the reference, whose
 * pipeline runs prym-rank on each finished WDM file, scripts/reproduce.py:
 * 263-266): every job submitted afterwards gets one pgw_rank_step over its
 * whole sequence, on the server's host workers with `threads` threads, after
 * its device memory is released.  keep_generator keeps the generator
 * matrices (terms x b x b values, about rank * b words). */
int pgw_server_enable_rank(pgw_server* s, uint32_t threads, uint32_t threshold,
                           uint32_t max_nlen_buffer, int keep_generator);
/* Rank-stage result of a finished job; call it before pgw_server_result,
 * which releases the job.  h_gen = NULL skips the generator copy. */
int pgw_server_rank(pgw_server* s, uint64_t ticket, pgw_rank_info* info, uint32_t* h_gen,
                    uint64_t gen_capacity, uint64_t* gen_terms, double* seconds);

/* ---- WDM output ------------------------------------------------------------ */

/* save_all_data + save_wdm_sym (cuda/cuprym_cyclic.cu:965-1007,
 * cuda/wdmfiles.h:164-222, zstd framing cuda/zstd_compat.h:210-333): writes
 * the symmetric WDM text of one sector via path + ".tmp" and a rename, as one
 * zstd frame (level 3) when path ends in .zst.  h_v, h_curv: cols x block
 * row-major; h_seq: length x block(block+1)/2 row-major.  The text is
 * formatted on `threads` host threads in ordered pieces and compressed by
 * libzstd's worker threads (threads = 0: PGW_WDM_THREADS or the host's
 * hardware threads, at most 16); the text is the same for every thread count.
 * Errors: PGW_ERR_IO. */
int pgw_wdm_write(const char* path, uint64_t rows, uint64_t cols, uint32_t prime,
                  const uint32_t* h_dr, const uint32_t* h_dc, const uint32_t* h_v,
                  const uint32_t* h_curv, uint32_t block, const uint32_t* h_seq, uint64_t length,
                  uint32_t threads);

/* Streaming form of pgw_wdm_write for a sequence that is still being
 * computed: the same file, byte for byte, with the text written while the
 * loop runs instead of after it (the reference formats and compresses on one
 * thread after the loop, cuda/cuprym_cyclic.cu:965-1007 called at :2461).
 * _open starts a writer thread that formats and compresses the header, Dr,
 * Dc and the V lines at once (h_dr, h_dc, h_v must stay valid until _finish
 * or _destroy).  _terms(h_seq, done): rows [0, done) of h_seq (length x
 * block(block+1)/2, the same array at every call) hold their final values;
 * the writer turns them into the entry lines' text in memory (up to
 * max_text_bytes of it; longer sequences are formatted at _finish).  The
 * progress callback of pgw_sequence_run is called exactly when this holds.
 * _finish(h_curv) writes curV and the entry lines, closes the zstd frame and
 * renames path.tmp to path.  _destroy of an unfinished stream stops the
 * writer and removes path.tmp.  Errors: PGW_ERR_IO. */
typedef struct pgw_wdm_stream pgw_wdm_stream;
int pgw_wdm_stream_open(const char* path, uint64_t rows, uint64_t cols, uint32_t prime,
                        const uint32_t* h_dr, const uint32_t* h_dc, const uint32_t* h_v,
                        uint32_t block, uint64_t length, uint32_t threads, uint64_t max_text_bytes,
                        pgw_wdm_stream** out);
int pgw_wdm_stream_terms(pgw_wdm_stream* s, const uint32_t* h_seq, uint64_t done);
int pgw_wdm_stream_finish(pgw_wdm_stream* s, const uint32_t* h_curv);
int pgw_wdm_stream_destroy(pgw_wdm_stream* s);



#ifdef __cplusplus
}
#endif
#endif

// pgw_kernels.h -- launchers of the gfx950 kernels (pgw_kernels.hip).
#pragma once

#include "pgw_internal.h"

namespace pgw {

constexpr size_t kMfmaZeroBytes = 1024;  // least zero region after the table digits

// Zero bytes after the n tables of an MFMA table: one whole table (so that
// table index n reads as an all-zero table) and at least 1 KB.
inline size_t mfma_zero_bytes(uint32_t nl, uint32_t rows_pad, uint32_t k_pad) {
    const size_t table = (size_t)nl * rows_pad * k_pad;
    return table > kMfmaZeroBytes ? table : kMfmaZeroBytes;
}

// Table-stationary digit planes of a synthetic multiplication table for the
// MFMA sweep: tabl[i][limb][R_pad][K_pad] int8 (balanced base-256 digits).
struct MfmaTable {
    int8_t* d = nullptr;
    const int8_t* zero = nullptr;  // the zero region after the digits (inside the allocation of d)
    uint32_t rows = 0, k_in = 0, rows_pad = 0, k_pad = 0;
};

struct ShiftWeights {
    uint32_t w[8];
};

// Kernel arguments of one MFMA sweep launch (pgw_mfma_sweep.h).
struct MfmaSweepArgs {
    const uint32_t* inc;    // incidence table of the sweep direction (n + 1 words per output panel)
    const int8_t* tabl;
    const int8_t* lp;
    const int8_t* zero;     // 1 KB of zero bytes (staging loads that have no data)
    uint32_t n, deg_out;
    uint64_t u_out, v_in;
    uint32_t r_out, r_pad, k_in, k_pad;
    uint32_t row_slices;
    uint64_t col_groups;
    uint64_t xcd_items;     // work items per XCD: the grid is 8 * xcd_items blocks
    uint32_t* out;
    const uint32_t* scale;
    ModP m;
    ShiftWeights sw;
};

// Launch description of one dense-table (r = 1) Koszul sweep.
struct DenseSweep {
    bool fwd = true;          // forward gather (out: (k-1)-subsets) or transpose (out: k-subsets)
    uint32_t n = 0, k = 0;    // exterior algebra Lambda^k of an n-dim space
    uint64_t u_out = 0;       // output subsets
    uint64_t v_in = 0;        // input subsets
    uint32_t r_out = 0;       // output rows per subset panel
    uint32_t k_in = 0;        // input rows per subset panel
    const uint64_t* binom = nullptr;  // device C(a,b) table, 64 x 64
    const uint32_t* inc = nullptr;    // MFMA path: incidence table (mfma_incidence_table)
};

// plain u64 path ------------------------------------------------------------
void launch_dense_plain(const DenseSweep& s, const uint32_t* tab, const uint32_t* tab_neg,
                        uint32_t k4, const uint32_t* in, uint32_t* out, const uint32_t* scale,
                        uint32_t b, const ModP& m, hipStream_t st);
// Sliced-ELL sparse map on the device (layout: SellHost, pgw_internal.h).
struct SellDevice {
    uint64_t rows = 0, cols = 0, nnz = 0, padded = 0;  // padded: stored entries
    uint64_t words = 0;                                 // u32 words of ent
    const uint64_t* slice = nullptr;
    // form 0: (column, coefficient) u32 pairs, slice[] in entries;
    // form 1: one u32 per entry, column in the low pack_bits bits and the
    //   coefficient above them (when the column count and p fit together in
    //   32 bits), slice[] in entries;
    // form 2: per slice, u32 words holding the column's offset from the
    //   slice's base column (sbase[s] = base | 1 << 31) when the slice's
    //   columns span less than 2^pack_bits, else pairs (sbase[s] = 0);
    //   slice[] in u32 words, every slice starting at an even word
    int form = 0;
    const uint32_t* ent = nullptr;
    const uint32_t* sbase = nullptr;
    uint32_t pack_bits = 0;
};
// out (M.rows x b) = M in (M.cols x b) mod p; out must not overlap in.  With
// out_row, row r of M lands at row out_row[r] of out (a permutation of the
// rows; a scattered 16-byte store per row instead of a coalesced one).
// With bump, the launch also adds 4 to *bump (the sequence loop's position).
void launch_sell_spmm(const SellDevice& M, const uint32_t* in, uint32_t* out, uint32_t b,
                      const ModP& m, hipStream_t st, const uint32_t* out_row = nullptr,
                      uint32_t* bump = nullptr);
// b = 4, p < 2^16: out = M in and, in the same launch, the sequence round's
// two packed projections of out with `other` (the apply's input block):
// mode 1 writes other^T out to out0 and out^T out to out1, mode 2 out^T other
// and out^T out (each at row *pos + pos_off when pos is given), through
// k_gram_finish over the launch's partials in `part`.  Returns false, having
// launched nothing, when the shape does not qualify (p, alignment, or more
// partials than part_bytes hold).  With nblocks, the finish is left to the
// caller (launch_gram_finish_round) and *nblocks receives the partial count.
bool launch_sell_spmm_gram(const SellDevice& M, const uint32_t* in, uint32_t* out, const uint32_t* other,
                           int mode, uint32_t* out0, uint32_t* out1, uint64_t* part, size_t part_bytes,
                           const uint32_t* pos, uint32_t pos_off, const ModP& m, hipStream_t st,
                           unsigned* nblocks = nullptr);
// Partial workgroups launch_sell_spmm_gram uses on M with part_bytes of
// scratch, or 0 when the shape does not qualify (pointer alignment aside).
unsigned sell_gram_plan(const SellDevice& M, size_t part_bytes, const ModP& m);
// Both halves of a b = 4 sequence round at once: the first half's partials
// (part_a, na workgroups; mode 1) to rows *pos and *pos + 1 of out0 / out1,
// the second's (part_b, nb; mode 2) to rows *pos + 2 and *pos + 3.
void launch_gram_finish_round(const uint64_t* part_a, unsigned na, const uint64_t* part_b, unsigned nb,
                              uint32_t* out0, uint32_t* out1, const ModP& m, const uint32_t* pos, hipStream_t st);

// elementwise -----------------------------------------------------------------
void launch_scale_rows(const uint32_t* in, const uint32_t* s, uint32_t* out, uint64_t rows,
                       uint32_t b, const ModP& m, hipStream_t st);
// The row-mixing chains of an operator on the device (MixChains, pgw_internal.h).
struct MixChainsDevice {
    uint64_t n = 0;  // positions, ghosts included
    const uint32_t* row = nullptr;
    const uint32_t* a = nullptr;
    const uint32_t* b = nullptr;
    const uint32_t* dr = nullptr;
    const uint8_t* flag = nullptr;
};
// out = P^T Dr P in for the one or two rounds the chains encode, in one pass
// (in and out distinct; every row of out written once).
void launch_mix_chains(const MixChainsDevice& c, const uint32_t* in, uint32_t* out, uint32_t b, const ModP& m,
                       hipStream_t st);
// P^T Dr P on a target block in place (P = `rounds` rounds of paired 2x2 maps).
// `unpaired_row` is the row missing from the last round's pairs (odd row count) or -1.
void launch_target_precond(uint32_t* block, const pgw_rowmix_pair* pairs, uint64_t pair_count,
                           uint32_t rounds, const uint32_t* dr, int64_t unpaired_row,
                           uint64_t rows, uint32_t b, const ModP& m, hipStream_t st);

// Gram projection -----------------------------------------------------------
// mode 0: out0 = X^T Y; mode 1: out0 = X^T Y, out1 = Y^T Y; mode 2: out0 = X^T Y, out1 = X^T X.
size_t gram_scratch_bytes(uint32_t b);
// pos (optional, device): outputs land ut * (*pos + pos_off) entries after out0/out1
// mfma: the partial sums on int8 MFMA (pgw_gram_mfma.hip, b in {4, 8, 16}),
// else on the vector ALUs; both write the same partials, the same results.
void launch_gram(const uint32_t* x, const uint32_t* y, uint64_t rows, uint32_t b, int mode,
                 uint32_t* out0, uint32_t* out1, uint64_t* scratch, const ModP& m,
                 hipStream_t st, const uint32_t* pos = nullptr, uint32_t pos_off = 0, bool mfma = false);
// MFMA partials (pgw_gram_mfma.hip): workgroups for `rows`, and the launch
unsigned gram_mfma_grid(uint64_t rows, uint32_t b, unsigned max_grid);
void launch_gram_mfma_partial(const uint32_t* x, const uint32_t* y, uint64_t rows, uint32_t b, int mode,
                              uint64_t* part, unsigned grid, const ModP& m, hipStream_t st);
// Allocation-free Gram (mode 0 only, out = X^T Y packed): partials in a
// fixed device array of gram_noalloc_scratch_bytes() (2 MiB) per GPU.
// gram_noalloc_setup resolves that array's address on `device` (call it with
// the device current, once, before the first launch on it).
size_t gram_noalloc_scratch_bytes();
void gram_noalloc_setup(int device);
void launch_gram_noalloc(int device, const uint32_t* x, const uint32_t* y, uint64_t rows,
                         uint32_t b, uint32_t* out, const ModP& m, hipStream_t st);
// *pos += by (one thread)
void launch_advance_pos(uint32_t* pos, uint32_t by, hipStream_t st);

// kernel-vector generator slice: D[r][c] += sum_i B[r][i] G[c*b+i]
void launch_accumulate_generator(const uint32_t* B, const uint32_t* G, uint32_t* D, uint64_t rows,
                                 uint32_t b, const ModP& m, hipStream_t st);

// int8-limb MFMA path -----------------------------------------------------------
uint32_t mfma_limbs(uint32_t p);  // 2 for p <= 65536, else 4
bool mfma_block_supported(uint32_t b);
// host: digits of a table for the MFMA sweep (rows x k_in view given by accessor)
void build_mfma_table(MfmaTable& t, uint32_t n, uint32_t rows, uint32_t k_in, uint32_t p,
                      const uint32_t* src, bool transpose_src);
// per-sweep digit-plane layout
uint64_t limb_plane_bytes(uint64_t v, uint32_t k_in, uint32_t b, uint32_t nl);
// Incidence table of one sweep direction for the MFMA kernels: n + 1 words
// per output panel u (a deg_out-subset of [0, n), lexicographic rank order):
// word 0 = mask of the tables incident to u (forward: i not in u; transpose:
// i in u), word 1 + i = rank of the neighbour subset (u with i added or
// removed) | sign << 31 (sign = parity of the elements of u below i), or ~0
// when u is not incident to table i.
std::vector<uint32_t> mfma_incidence_table(uint32_t n, uint32_t deg_out, bool fwd, uint64_t u_out,
                                           uint64_t v_in);
uint32_t limb_stride(uint32_t k_in);
void launch_limb_split(const uint32_t* in, const uint32_t* scale, int8_t* lp, uint64_t v,
                       uint32_t k_in, uint32_t b, uint32_t nl, const ModP& m, hipStream_t st);
// Grouped launches (pgw_batch.cpp): the same stage of many jobs in one grid;
// job j owns blocks [first[j], first[j+1]) of it, its arguments at jobs[j]
// (both arrays in device memory).
// The job of this block in a grouped launch (uniform across the block):
// first[lo] <= blockIdx.x < first[lo + 1].  Ranges are multiples of 8 blocks
// for the MFMA sweeps, so that a block's XCD (hardware block id mod 8) is its
// local id mod 8.
// Grid of a persistent launch on the current device: per_cu workgroups per
// CU, a multiple of 8 (the XCD count).
uint32_t persistent_grid(uint32_t per_cu);

// the job that owns block `blk` of a grouped grid (first[j] <= blk < first[j + 1])
__device__ __forceinline__ uint32_t group_job_at(const uint32_t* __restrict__ first, uint32_t njobs,
                                                 uint32_t blk) {
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (first[mid] <= blk) lo = mid; else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ uint32_t group_job(const uint32_t* __restrict__ first, uint32_t njobs) {
    return group_job_at(first, njobs, blockIdx.x);
}

struct LimbJob {
    const uint32_t* in;
    const uint32_t* scale;
    int8_t* lp;
    uint64_t v_count;
    uint32_t k_in, k_stride;
    ModP m;
};
uint64_t limb_group_blocks(const LimbJob& j, uint32_t b);
void launch_limb_split_group(const LimbJob* jobs, const uint32_t* first, uint32_t njobs,
                             uint64_t blocks, uint32_t b, uint32_t nl, hipStream_t st);
struct MixJob {
    const pgw_rowmix_pair* pairs;  // round r at pairs + r * count
    uint64_t count;
    uint32_t* block;
    const uint32_t* dr;
    int64_t unpaired;              // row outside the last round's pairs, or -1
    ModP m;
};
uint64_t mix_group_blocks(const MixJob& j, uint32_t b);
// mode 0 forward round, 1 transposed round, 2 the fused P_r^T Dr P_r (plus
// the unpaired row's Dr); m31: every job's p is 2^31 - 1
void launch_rowmix_group(const MixJob* jobs, const uint32_t* first, uint32_t njobs, uint64_t blocks,
                         uint32_t b, int mode, bool m31, uint32_t round, hipStream_t st);
struct GramJob {
    const uint32_t* x;
    const uint32_t* y;
    uint64_t rows;
    uint64_t* part;    // the job's Gram scratch
    uint32_t nblocks;  // its partial workgroups (gram_group_blocks)
    uint32_t* out0;
    uint32_t* out1;
    const uint32_t* pos;
    ModP m;
};
uint32_t gram_group_blocks(uint64_t rows);
// partial pass over blocks [first_part[j], ...), finish pass over
// [first_fin[j], ...) (P or 2P entries per job)
void launch_gram_group(const GramJob* jobs, const uint32_t* first_part, const uint32_t* first_fin,
                       uint32_t njobs, uint64_t part_blocks, uint64_t fin_blocks, uint32_t b, int mode,
                       uint32_t pos_off, hipStream_t st);
// *pos[j] += by for every job
void launch_advance_pos_group(uint32_t* const* pos, uint32_t njobs, uint32_t by, hipStream_t st);

// Grouped sweeps (pgw_batch.cpp): many jobs of one block width, limb count,
// direction and register-A tile configuration (7-9) in one launch.  prepare
// fills a job's arguments and returns its block count (a multiple of 8);
// launch runs jobs[0, njobs) whose blocks start at first[j] (first[njobs] =
// blocks), both arrays in device memory.
uint64_t mfma_group_prepare(int cfg, uint32_t b, uint32_t nl, const DenseSweep& s, const MfmaTable& t,
                            const int8_t* lp, uint32_t* out, const uint32_t* scale, const ModP& m,
                            MfmaSweepArgs& a);
void mfma_group_launch(int cfg, uint32_t b, uint32_t nl, bool fwd, const MfmaSweepArgs* jobs,
                       const uint32_t* first, uint32_t njobs, uint64_t blocks, bool persistent, hipStream_t st);
// cfg selects a tile configuration in [0, mfma_config_count())
int mfma_config_count();
void launch_dense_mfma(const DenseSweep& s, const MfmaTable& t, const int8_t* lp, uint32_t* out,
                       const uint32_t* scale, uint32_t b, uint32_t nl, const ModP& m,
                       hipStream_t st, int cfg);

}  // namespace pgw

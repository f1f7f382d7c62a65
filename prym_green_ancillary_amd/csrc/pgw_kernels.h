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
    const uint64_t* slice = nullptr;
    const uint32_t* ent = nullptr;
};
// out (M.rows x b) = M in (M.cols x b) mod p; out must not overlap in
void launch_sell_spmm(const SellDevice& M, const uint32_t* in, uint32_t* out, uint32_t b,
                      const ModP& m, hipStream_t st);

// elementwise -----------------------------------------------------------------
void launch_scale_rows(const uint32_t* in, const uint32_t* s, uint32_t* out, uint64_t rows,
                       uint32_t b, const ModP& m, hipStream_t st);
// P^T Dr P on a target block in place (P = `rounds` rounds of paired 2x2 maps).
// `unpaired_row` is the row missing from the last round's pairs (odd row count) or -1.
void launch_target_precond(uint32_t* block, const pgw_rowmix_pair* pairs, uint64_t pair_count,
                           uint32_t rounds, const uint32_t* dr, int64_t unpaired_row,
                           uint64_t rows, uint32_t b, const ModP& m, hipStream_t st);

// Gram projection -----------------------------------------------------------
// mode 0: out0 = X^T Y; mode 1: out0 = X^T Y, out1 = Y^T Y; mode 2: out0 = X^T Y, out1 = X^T X.
size_t gram_scratch_bytes(uint32_t b);
// pos (optional, device): outputs land ut * (*pos + pos_off) entries after out0/out1
void launch_gram(const uint32_t* x, const uint32_t* y, uint64_t rows, uint32_t b, int mode,
                 uint32_t* out0, uint32_t* out1, uint64_t* scratch, const ModP& m,
                 hipStream_t st, const uint32_t* pos = nullptr, uint32_t pos_off = 0);
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
// cfg selects a tile configuration in [0, mfma_config_count())
int mfma_config_count();
void launch_dense_mfma(const DenseSweep& s, const MfmaTable& t, const int8_t* lp, uint32_t* out,
                       const uint32_t* scale, uint32_t b, uint32_t nl, const ModP& m,
                       hipStream_t st, int cfg);

}  // namespace pgw

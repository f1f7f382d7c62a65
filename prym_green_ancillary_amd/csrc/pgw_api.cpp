// pgw_api.cpp -- the C ABI of libpgw: contexts, operators, the block-apply
// chain, the Gram projection, the sequence loop and kernel-vector recovery.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "pgw_op.h"

using namespace pgw;

namespace {

uint32_t bits_for(uint64_t v) {  // bits of the largest value below v (at least 1)
    uint32_t b = 1;
    while (b < 64 && (v - 1) >> b) ++b;
    return b;
}

// A sparse map in its device (sliced-ELL) form, owned by the operator's arena.
// When a column index and a coefficient (< p) fit in 32 bits together, every
// entry is one packed u32 (column | coefficient << column bits): half the
// bytes the map streams per apply (the paper instances up to g26).  When they
// do not, each slice of kSellHeight rows whose columns span less than
// 2^(32 - bits(p)) stores its entries as packed offsets from its smallest
// column, the others as pairs (g28: all but 177 of M1's and M1^T's 148 k
// slices).  PGW_SELL_PACK=0 keeps pairs everywhere and PGW_SELL_PACK=slice
// forces the per-slice form (test knobs); PGW_SELL_SLICE_BITS narrows the
// per-slice offsets so that a test meets both kinds of slice.
SellDevice upload_sell(DeviceArena& A, const CsrHost& csr, uint32_t p) {
    const SellHost h = csr_to_sell(csr);
    SellDevice d;
    d.rows = h.rows;
    d.cols = h.cols;
    d.nnz = h.nnz;
    d.padded = h.slice.back();
    const uint32_t cb = bits_for(std::max<uint64_t>(h.cols, 2));
    const char* knob = std::getenv("PGW_SELL_PACK");
    const bool pairs_only = knob && knob[0] == '0';
    const bool force_slice = knob && std::strcmp(knob, "slice") == 0;
    if (!pairs_only && !force_slice && cb + bits_for(p) <= 32) {
        std::vector<uint32_t> packed(d.padded);
        for (uint64_t q = 0; q < d.padded; ++q) {
            const uint32_t col = h.ent[2 * q], coef = h.ent[2 * q + 1];
            require((uint64_t)col < (1ULL << cb) && (uint64_t)coef < (1ULL << (32 - cb)),
                    "sparse map entry outside its packed range");
            packed[q] = col | (coef << cb);
        }
        d.slice = A.upload(h.slice);
        d.ent = A.upload(packed);
        d.words = d.padded;
        d.form = 1;
        d.pack_bits = cb;
        return d;
    }
    uint32_t ob = 32 - bits_for(p);  // offset bits of a packed slice
    const char* narrow = std::getenv("PGW_SELL_SLICE_BITS");
    if (narrow) ob = std::min<uint32_t>(ob, (uint32_t)std::max(0, std::atoi(narrow)));
    if (pairs_only || (!narrow && ob < 8) || ob < 2) {
        d.slice = A.upload(h.slice);
        d.ent = A.upload(h.ent);
        d.words = 2 * d.padded;
        return d;
    }
    const uint64_t nslices = h.slice.size() - 1;
    std::vector<uint64_t> wslice(nslices + 1, 0);
    std::vector<uint32_t> sbase(nslices, 0);
    for (uint64_t s = 0; s < nslices; ++s) {
        uint32_t lo = ~0u, hi = 0;
        for (uint64_t q = h.slice[s]; q < h.slice[s + 1]; ++q)
            if (h.ent[2 * q + 1] != 0) {  // padding entries carry coefficient 0
                lo = std::min(lo, h.ent[2 * q]);
                hi = std::max(hi, h.ent[2 * q]);
            }
        const uint64_t n = h.slice[s + 1] - h.slice[s];
        const bool packed = lo == ~0u || (uint64_t)(hi - lo) < (1ULL << ob);
        if (packed) sbase[s] = (lo == ~0u ? 0u : lo) | 0x80000000u;
        // pairs start at an even word, so the kernel reads them as 8-byte loads
        const uint64_t at = packed ? wslice[s] : (wslice[s] + 1) / 2 * 2;
        wslice[s] = at;
        wslice[s + 1] = at + (packed ? n : 2 * n);
    }
    std::vector<uint32_t> words(wslice[nslices], 0);
    parallel_rows(nslices, [&](uint64_t lo_s, uint64_t hi_s) {
        for (uint64_t s = lo_s; s < hi_s; ++s) {
            const bool packed = sbase[s] >> 31;
            const uint32_t base = sbase[s] & 0x7fffffffu;
            for (uint64_t q = h.slice[s]; q < h.slice[s + 1]; ++q) {
                const uint64_t i = q - h.slice[s];
                const uint32_t col = h.ent[2 * q], coef = h.ent[2 * q + 1];
                if (packed) {
                    require(coef == 0 || col - base < (1u << ob), "sparse map entry outside its slice range");
                    words[wslice[s] + i] = coef == 0 ? 0u : (col - base) | (coef << ob);
                } else {
                    words[wslice[s] + 2 * i] = col;
                    words[wslice[s] + 2 * i + 1] = coef;
                }
            }
        }
    });
    require(d.cols < (1ULL << 31), "per-slice packing needs fewer than 2^31 columns");
    d.slice = A.upload(wslice);
    d.ent = A.upload(words);
    d.sbase = A.upload(sbase);
    d.words = words.size();
    d.form = 2;
    d.pack_bits = ob;
    return d;
}

}  // namespace

// ---------------------------------------------------------------------------

namespace {

DenseSweep sweep_desc(const pgw_op* op, int which) {
    DenseSweep s;
    s.n = op->n;
    s.k = op->k;
    s.fwd = which == 0 || which == 2;
    s.u_out = s.fwd ? op->u_tgt : op->u_src;
    s.v_in = s.fwd ? op->u_src : op->u_tgt;
    s.r_out = s.fwd ? op->Mp : op->M;
    s.k_in = s.fwd ? op->M : op->Mp;
    s.binom = op->ctx->d_binom;
    s.inc = op->d_inc[s.fwd ? 0 : 1];
    return s;
}

void check_mfma_bound(const DenseSweep& s, uint32_t nl) {
    // int32 accumulators: incidences * K_in * (pairs per shift <= NL) * 2^14 < 2^31
    const uint64_t inc = s.fwd ? (s.n - s.k + 1) : s.k;
    require(inc * s.k_in * nl * (1ULL << 14) < (1ULL << 31),
            "int8-limb accumulation bound exceeded for this operator (incidences*K*limbs >= 2^17)",
            PGW_ERR_UNSUPPORTED);
}

// Events around one stage launch while op->profile is set (benchmarks only).
struct StageTimer {
    pgw_op* op;
    hipStream_t st;
    int stage;
    double bytes;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    StageTimer(pgw_op* o, int stg, double by, hipStream_t s) : op(o), st(s), stage(stg), bytes(by) {
        if (op->profile) {
            PGW_HIP(hipEventCreate(&e0));
            PGW_HIP(hipEventCreate(&e1));
            PGW_HIP(hipEventRecord(e0, st));
        }
    }
    void done() {
        if (op->profile) {
            PGW_HIP(hipEventRecord(e1, st));
            op->stage_events.push_back({stage, bytes, e0, e1});
        }
    }
};

// algorithmic bytes of the elementwise stages (each stage's unique input and
// output bytes, the quantities an HBM roofline prices)
double limb_split_bytes(uint64_t v, uint32_t k_in, uint32_t b, uint32_t nl, bool scaled) {
    const double elems = (double)v * k_in;
    return elems * b * 4 + (scaled ? elems * 4 : 0) + 2.0 * nl * v * b * limb_stride(k_in);
}
double rowmix_bytes(const pgw_op* op) {
    const double rb = (double)op->rows * op->b * 4;
    if (op->rounds == 0 || op->pair_count == 0) return 2 * rb + op->rows * 4.0;
    // 2R-1 launches, each reading and writing the paired rows and its pair records
    return (2.0 * op->rounds - 1) * (2 * rb + 16.0 * op->pair_count) + op->rows * 4.0;
}
double sell_bytes(const SellDevice& M, uint32_t b) {
    const uint64_t slices = (M.rows + kSellHeight - 1) / kSellHeight;
    return (double)M.words * 4 + (slices + 1) * (M.form == 2 ? 12.0 : 8.0) + (double)M.cols * b * 4 +
           (double)M.rows * b * 4;
}

void timed_limb_split(pgw_op* op, const uint32_t* in, const uint32_t* scale, int8_t* lp, uint64_t v,
                      uint32_t k_in, hipStream_t st) {
    StageTimer t(op, 1, limb_split_bytes(v, k_in, op->b, op->nl, scale != nullptr), st);
    launch_limb_split(in, scale, lp, v, k_in, op->b, op->nl, op->m, st);
    t.done();
}

void timed_sell(pgw_op* op, const SellDevice& M, const uint32_t* in, uint32_t* out, hipStream_t st,
                const uint32_t* out_row = nullptr, uint32_t* bump = nullptr) {
    StageTimer t(op, 4, sell_bytes(M, op->b) + (out_row ? M.rows * 4.0 : 0.0), st);
    launch_sell_spmm(M, in, out, op->b, op->m, st, out_row, bump);
    t.done();
}

// in: u32 block (plain) ; the MFMA path consumes the digit planes prepared by
// the caller in op->d_lp_src / d_lp_mid
void run_sweep(pgw_op* op, int which, const uint32_t* in_u32, const int8_t* in_lp, uint32_t* out,
               const uint32_t* scale, hipStream_t st) {
    const DenseSweep s = sweep_desc(op, which);
    const double in_bytes = op->method == PGW_METHOD_MFMA
                                ? 2.0 * op->nl * s.v_in * op->b * limb_stride(s.k_in)
                                : (double)s.v_in * s.k_in * op->b * 4;
    StageTimer t(op, 0, in_bytes + (double)s.u_out * s.r_out * op->b * 4, st);
    if (op->method == PGW_METHOD_MFMA)
        launch_dense_mfma(s, op->mt[which], in_lp, out, scale, op->b, op->nl, op->m, st,
                          op->mfma_cfg[which]);
    else
        launch_dense_plain(s, op->d_tab[which][0], op->d_tab[which][1], op->k4[which], in_u32, out,
                           scale, op->b, op->m, st);
    t.done();
}

// F * (Dc x) into op->d_target (rows x b)
void forward_into_target(pgw_op* op, const uint32_t* x, hipStream_t st) {
    const ModP& m = op->m;
    if (op->kind == PGW_KIND_SYNTHETIC) {
        if (op->method == PGW_METHOD_MFMA) {
            timed_limb_split(op, x, op->d_dc, op->d_lp_src, op->u_src, op->M, st);
            run_sweep(op, 0, nullptr, op->d_lp_src, op->d_mid, nullptr, st);
            timed_limb_split(op, op->d_mid, nullptr, op->d_lp_mid, op->u_tgt, op->Mp, st);
            run_sweep(op, 1, nullptr, op->d_lp_mid, op->d_target, nullptr, st);
        } else {
            {
                StageTimer t(op, 3, 2.0 * op->cols * op->b * 4 + op->cols * 4.0, st);
                launch_scale_rows(x, op->d_dc, op->d_scaled, op->cols, op->b, m, st);
                t.done();
            }
            run_sweep(op, 0, op->d_scaled, nullptr, op->d_mid, nullptr, st);
            run_sweep(op, 1, op->d_mid, nullptr, op->d_target, nullptr, st);
        }
    } else {
        // F (Dc x) of the cyclic operator: the first two sparse maps
        timed_sell(op, op->m1, x, op->d_y, st);
        timed_sell(op, op->m2, op->d_y, op->d_target, st);
    }
}

// out = Dc F^T (target), synthetic operator
void backward_from_target(pgw_op* op, const uint32_t* target, uint32_t* out, hipStream_t st) {
    if (op->method == PGW_METHOD_MFMA) {
        timed_limb_split(op, target, nullptr, op->d_lp_src, op->u_src, op->M, st);
        run_sweep(op, 2, nullptr, op->d_lp_src, op->d_mid, nullptr, st);
        timed_limb_split(op, op->d_mid, nullptr, op->d_lp_mid, op->u_tgt, op->Mp, st);
        run_sweep(op, 3, nullptr, op->d_lp_mid, out, op->d_dc, st);
    } else {
        run_sweep(op, 2, target, nullptr, op->d_mid, nullptr, st);
        run_sweep(op, 3, op->d_mid, nullptr, out, op->d_dc, st);
    }
}


}  // namespace

namespace {

// S = M1^T M2^T M3 M2 M1 up to the last map: Dc, the elimination and
// P^T Dr P are folded in.  With the row-mixing order, M2 stores t by
// position, M3 (kept in that order) gathers from nearby positions and stores
// its rows back in natural order, where M2^T gathers them.  Leaves M2^T M3 M2
// M1 in in op->d_y; with bump, M1's launch also adds 4 to *bump.
void cyclic_front(pgw_op* op, const uint32_t* in, hipStream_t st, uint32_t* bump = nullptr) {
    timed_sell(op, op->m1, in, op->d_y, st, nullptr, bump);
    timed_sell(op, op->m2, op->d_y, op->d_target, st, op->d_t_pos);
    timed_sell(op, op->m3, op->d_target, op->d_t2, st, op->d_t_row);
    timed_sell(op, op->m2t, op->d_t2, op->d_y, st);
}

// sequence loops of at least this many rounds prepare the row-mixing chains
constexpr uint64_t kPrepareRounds = 1024;

// PGW_CYCLIC_FUSED_GRAM=0 keeps the projection in its own launches (a test
// knob, read when a round is enqueued or captured)
bool fused_gram_allowed() {
    const char* knob = std::getenv("PGW_CYCLIC_FUSED_GRAM");
    return !(knob && knob[0] == '0');
}

// Whether a sequence round of `op` runs as ten sparse maps (the two last
// with the projections folded in, each half's partials in its own half of
// the operator's Gram scratch) and one finish for both halves, the position
// moving on in the round's first map; *nb receives the partial count.
bool fused_round_plan(const pgw_op* op, unsigned* nb) {
    if (op->kind == PGW_KIND_SYNTHETIC || op->b != 4 || op->gram_mfma || !fused_gram_allowed() || op->m1.rows == 0)
        return false;
    *nb = sell_gram_plan(op->m1t, gram_scratch_bytes(op->b) / 2, op->m);
    return *nb != 0;
}

}  // namespace

void pgw::apply_square_impl(pgw_op* op, const uint32_t* in, uint32_t* out, hipStream_t st) {
    if (op->kind != PGW_KIND_SYNTHETIC) {
        cyclic_front(op, in, st);
        timed_sell(op, op->m1t, op->d_y, out, st);
        return;
    }
    forward_into_target(op, in, st);
    if (op->chains.n) {
        // P^T Dr P in one pass along the row-mixing chains, into d_target2
        StageTimer t(op, 2, 2.0 * op->rows * op->b * 4 + op->chains.n * 17.0, st);
        launch_mix_chains(op->chains, op->d_target, op->d_target2, op->b, op->m, st);
        t.done();
        backward_from_target(op, op->d_target2, out, st);
        return;
    }
    {
        StageTimer t(op, 2, rowmix_bytes(op), st);
        launch_target_precond(op->d_target, op->d_pairs, op->pair_count, op->rounds, op->d_dr,
                              op->unpaired, op->rows, op->b, op->m, st);
        t.done();
    }
    backward_from_target(op, op->d_target, out, st);
}

void pgw::apply_and_project(pgw_op* op, const uint32_t* in, uint32_t* out, int mode, uint32_t* seq0,
                            uint32_t* seq1, const uint32_t* pos, uint32_t pos_off, hipStream_t st) {
    if (op->kind != PGW_KIND_SYNTHETIC && op->b == 4 && !op->gram_mfma && fused_gram_allowed()) {
        cyclic_front(op, in, st);
        if (launch_sell_spmm_gram(op->m1t, op->d_y, out, in, mode, seq0, seq1, op->d_gram,
                                  gram_scratch_bytes(op->b), pos, pos_off, op->m, st))
            return;
        timed_sell(op, op->m1t, op->d_y, out, st);
    } else {
        apply_square_impl(op, in, out, st);
    }
    const uint32_t* x = mode == 1 ? in : out;
    const uint32_t* y = mode == 1 ? out : in;
    launch_gram(x, y, op->cols, op->b, mode, seq0, seq1, op->d_gram, op->m, st, pos, pos_off, op->gram_mfma);
}

namespace {

hipStream_t pick_stream(pgw_op* op, void* stream) {
    return stream ? (hipStream_t)stream : op->ctx->stream;
}

}  // namespace

namespace pgw {

void validate_desc_common(const pgw_operator_desc* d) {
    require(d != nullptr, "null operator description");
    require(d->prime > 2 && d->prime < (1u << 31), "prime must satisfy 2 < p < 2^31");
    for (uint32_t q = 2; (uint64_t)q * q <= d->prime; ++q)
        require(d->prime % q != 0, "modulus " + std::to_string(d->prime) + " is not prime");
    require(d->block >= 1 && d->block <= 16, "block width must be in 1..16");
}

}  // namespace pgw

namespace {

void synthetic_dims(const pgw_operator_desc* d, uint64_t& rows, uint64_t& cols, uint64_t& mid) {
    require(d->n >= 2 && d->n < 64, "synthetic operator needs 2 <= n < 64");
    require(d->k >= 1 && d->k <= d->n, "synthetic operator needs 1 <= k <= n");
    require(d->m_src >= 1 && d->m_tgt >= 1, "coefficient dimensions must be positive");
    const uint64_t us = host_binom(d->n, d->k), ut = host_binom(d->n, d->k - 1);
    require(us < (1ULL << 31) && ut < (1ULL << 31), "subset count too large");
    cols = us * d->m_src;
    rows = cols;
    mid = ut * d->m_tgt;
}

}  // namespace

namespace pgw {

bool is_deformation(uint32_t kind) {
    return kind == PGW_KIND_CYCLIC_AUGMENTED || kind == PGW_KIND_CYCLIC_REPLACEMENT;
}

// Source columns of a deformation variant over an eliminated base with
// `base` columns, after the checks of run_deformation_eliminated_variant
// (cuda/cuprym_deformation.cu:2600-2623, load_drop_columns_file :452-472):
// two dense columns; for the replacement two distinct base columns, which
// are returned ascending.
uint64_t deformation_cols(const pgw_operator_desc* d, uint64_t base, uint32_t& drop0, uint32_t& drop1) {
    require(d->dense_cols == 2, "deformation operators take a dense block of two columns");
    require(d->dense_block != nullptr, "deformation operator needs its dense block");
    drop0 = drop1 = 0;
    if (d->kind == PGW_KIND_CYCLIC_AUGMENTED) return base + d->dense_cols;
    drop0 = std::min(d->drop_cols[0], d->drop_cols[1]);
    drop1 = std::max(d->drop_cols[0], d->drop_cols[1]);
    require(drop1 < base, "drop-column index is outside the base column range");
    require(drop0 != drop1, "drop-column indices must be distinct");
    return base - 2 + d->dense_cols;
}

int64_t validate_rowmix_plan(uint64_t rows, uint32_t rounds, const pgw_rowmix_pair* pairs, uint32_t p) {
    const uint64_t pc = rows / 2;
    if (rounds == 0 || pc == 0) return -1;
    require(pairs != nullptr, "row-mix plan is required when rounds > 0");
    std::vector<uint8_t> seen(rows);
    for (uint32_t r = 0; r < rounds; ++r) {
        std::fill(seen.begin(), seen.end(), 0);
        for (uint64_t q = 0; q < pc; ++q) {
            const pgw_rowmix_pair& pr = pairs[(uint64_t)r * pc + q];
            require(pr.first < rows && pr.second < rows && pr.first != pr.second,
                    "row-mix pair outside the target rows");
            require(!seen[pr.first] && !seen[pr.second], "row-mix pairs of a round must be disjoint");
            require(pr.a >= 1 && pr.a < p && pr.b >= 1 && pr.b < p,
                    "row-mix coefficients must be nonzero residues");
            seen[pr.first] = seen[pr.second] = 1;
        }
    }
    int64_t unpaired = -1;
    if (rows % 2 == 1)
        for (uint64_t q = 0; q < rows; ++q)
            if (!seen[q]) unpaired = (int64_t)q;  // left out of the last round
    return unpaired;
}

void validate_preconditioners(const uint32_t* dr, uint64_t rows, const uint32_t* dc, uint64_t cols,
                              uint32_t p) {
    require(dr && dc, "preconditioners are required");
    uint32_t lo, hi;
    min_max_u32(dr, rows, lo, hi);
    require(rows == 0 || (lo >= 1 && hi < p), "row preconditioner entries must be nonzero residues");
    min_max_u32(dc, cols, lo, hi);
    require(cols == 0 || (lo >= 1 && hi < p), "column preconditioner entries must be nonzero residues");
}

}  // namespace pgw

// ---------------------------------------------------------------------------

namespace pgw {

void autotune_sweeps(pgw_op* op, uint32_t reps, int cfg_lo, int cfg_hi, uint32_t* h_choice, double* h_ms) {
    PGW_HIP(hipSetDevice(op->ctx->device));
    hipStream_t st = op->ctx->stream;
    if (op->method != PGW_METHOD_MFMA) {
        for (int w = 0; w < 4; ++w) {
            if (h_choice) h_choice[w] = 0;
            if (h_ms) h_ms[w] = 0.0;
        }
        return;
    }
    if (reps == 0) reps = 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PGW_HIP(hipEventCreate(&e0));
    try {
        PGW_HIP(hipEventCreate(&e1));
        // the sweeps' schedules do not depend on the data: time them on the
        // operator's own digit planes and scratch blocks
        for (int w = 0; w < 4; ++w) {
            const DenseSweep s = sweep_desc(op, w);
            const int8_t* lp = s.fwd ? op->d_lp_src : op->d_lp_mid;
            uint32_t* out = s.fwd ? op->d_mid : op->d_target;
            double best = 1e30;
            int best_cfg = op->mfma_cfg[w];
            for (int c = cfg_lo; c <= cfg_hi; ++c) {
                launch_dense_mfma(s, op->mt[w], lp, out, nullptr, op->b, op->nl, op->m, st, c);  // warm
                PGW_HIP(hipEventRecord(e0, st));
                for (uint32_t r = 0; r < reps; ++r)
                    launch_dense_mfma(s, op->mt[w], lp, out, nullptr, op->b, op->nl, op->m, st, c);
                PGW_HIP(hipEventRecord(e1, st));
                PGW_HIP(hipEventSynchronize(e1));
                float ms = 0;
                PGW_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms / reps < best) {
                    best = ms / reps;
                    best_cfg = c;
                }
            }
            op->mfma_cfg[w] = best_cfg;
            if (h_choice) h_choice[w] = (uint32_t)best_cfg;
            if (h_ms) h_ms[w] = best;
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

}  // namespace pgw

extern "C" {

int pgw_ctx_create(int device, pgw_ctx** out) {
    return guarded([&] {
        require(out != nullptr, "null output pointer");
        int count = 0;
        PGW_HIP(hipGetDeviceCount(&count));
        require(device >= 0 && device < count, "device index out of range");
        PGW_HIP(hipSetDevice(device));
        auto* c = new pgw_ctx;
        c->device = device;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            throw Error(PGW_ERR_HIP, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
        }
        upload_binomials(device);
        gram_noalloc_setup(device);
        std::vector<uint64_t> binom(64 * 64, 0);
        for (uint32_t a = 0; a < 64; ++a)
            for (uint32_t b = 0; b <= a; ++b) {
                unsigned __int128 v = 1;
                const uint32_t kk = std::min(b, a - b);
                for (uint32_t q = 0; q < kk; ++q) v = v * (a - q) / (q + 1);
                binom[a * 64 + b] = v > (unsigned __int128)UINT64_MAX ? UINT64_MAX : (uint64_t)v;
            }
        e = hipMalloc(&c->d_binom, binom.size() * sizeof(uint64_t));
        if (e == hipSuccess)
            e = hipMemcpy(c->d_binom, binom.data(), binom.size() * sizeof(uint64_t), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipStreamDestroy(c->stream);
            delete c;
            throw Error(PGW_ERR_HIP, std::string("binomial table upload failed: ") + hipGetErrorString(e));
        }
        *out = c;
    });
}

int pgw_ctx_destroy(pgw_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return;
        const uint64_t live = ctx->live.load();
        require(live == 0, "context still owns " + std::to_string(live) +
                               " operators or batches: destroy them first");
        PGW_HIP(hipSetDevice(ctx->device));
        PGW_HIP(hipStreamSynchronize(ctx->stream));
        PGW_HIP(hipStreamDestroy(ctx->stream));
        if (ctx->d_binom) (void)hipFree(ctx->d_binom);
        delete ctx;
    });
}

int pgw_ctx_synchronize(pgw_ctx* ctx) {
    return guarded([&] {
        require(ctx != nullptr, "null context");
        PGW_HIP(hipSetDevice(ctx->device));
        PGW_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int pgw_ctx_device_info(pgw_ctx* ctx, char* name, size_t name_len, uint64_t* total_bytes) {
    return guarded([&] {
        require(ctx != nullptr, "null context");
        hipDeviceProp_t prop;
        PGW_HIP(hipGetDeviceProperties(&prop, ctx->device));
        if (name && name_len) {
            // the marketing name can be empty where the driver's id table is missing
            std::snprintf(name, name_len, "%s%s%s, %d CUs", prop.name, prop.name[0] ? " " : "",
                          prop.gcnArchName, prop.multiProcessorCount);
        }
        if (total_bytes) *total_bytes = prop.totalGlobalMem;
    });
}

int pgw_operator_dims(const pgw_operator_desc* desc, uint64_t* rows, uint64_t* cols) {
    return guarded([&] {
        validate_desc_common(desc);
        uint64_t r = 0, c = 0, mid = 0;
        if (desc->kind == PGW_KIND_SYNTHETIC) {
            synthetic_dims(desc, r, c, mid);
        } else if (desc->kind == PGW_KIND_CYCLIC_ELIMINATED || is_deformation(desc->kind)) {
            CyclicLayout L;
            build_cyclic_layout(L, *desc);
            r = L.rows;
            c = L.cols;
            uint32_t d0, d1;
            if (is_deformation(desc->kind)) c = deformation_cols(desc, c, d0, d1);
        } else {
            throw Error(PGW_ERR_INVALID, "unknown operator kind");
        }
        if (rows) *rows = r;
        if (cols) *cols = c;
    });
}

int pgw_operator_create(pgw_ctx* ctx, const pgw_operator_desc* desc,
                        const uint32_t* h_row_precond, const uint32_t* h_col_precond,
                        uint32_t rowmix_rounds, const pgw_rowmix_pair* h_pairs, pgw_op** out) {
    pgw_op* op = nullptr;
    int rc = guarded([&] {
        require(ctx != nullptr && out != nullptr, "null context or output pointer");
        validate_desc_common(desc);
        require(h_row_precond && h_col_precond, "preconditioners are required");
        PGW_HIP(hipSetDevice(ctx->device));
        op = new pgw_op;
        op->arena.lock = &ctx->capture_mtx;
        op->ctx = ctx;
        op->kind = desc->kind;
        op->p = desc->prime;
        op->b = desc->block;
        op->m = make_modp(op->p);
        DeviceArena& A = op->arena;
        const uint32_t b = op->b;
        CyclicLayout layout;
        bool cyclic = false;
        if (desc->kind == PGW_KIND_SYNTHETIC) {
            require(desc->table_d && desc->table_e, "synthetic operator needs both tables");
            synthetic_dims(desc, op->rows, op->cols, op->mid);
            op->n = desc->n;
            op->k = desc->k;
            op->M = desc->m_src;
            op->Mp = desc->m_tgt;
            op->u_src = host_binom(op->n, op->k);
            op->u_tgt = host_binom(op->n, op->k - 1);
            const uint64_t tsize = (uint64_t)op->n * op->Mp * op->M;
            for (uint64_t q = 0; q < tsize; ++q)
                require(desc->table_d[q] < op->p && desc->table_e[q] < op->p,
                        "table entries must be reduced modulo p");
            uint32_t method = desc->method;
            const bool mfma_ok = mfma_block_supported(b) && op->n <= 32 && op->k >= 1 && op->k < 16;
            if (method == PGW_METHOD_AUTO) method = mfma_ok ? PGW_METHOD_MFMA : PGW_METHOD_PLAIN;
            require(method == PGW_METHOD_PLAIN || mfma_ok,
                    "MFMA path needs b in {4, 8, 16}, n <= 32 and k < 16", PGW_ERR_UNSUPPORTED);
            op->method = method;
            op->gram_mfma = method == PGW_METHOD_MFMA;
            // sweep tables: 0 = D (T, forward), 1 = E (E^T, transpose), 2 = E^T (E, forward),
            // 3 = D^T (T^T, transpose)
            const uint32_t* src[4] = {desc->table_d, desc->table_e, desc->table_e, desc->table_d};
            if (method == PGW_METHOD_MFMA) {
                op->nl = mfma_limbs(op->p);
                if (const char* env = std::getenv("PGW_MFMA_CFG")) {
                    const int c = std::atoi(env);
                    require(c >= 0 && c < mfma_config_count(), "PGW_MFMA_CFG out of range");
                    for (int& x : op->mfma_cfg) x = c;
                }
                for (int w = 0; w < 2; ++w) {  // sweep 0 is forward, sweep 1 transpose
                    const DenseSweep s = sweep_desc(op, w);
                    op->d_inc[w] = A.upload(mfma_incidence_table(op->n, s.fwd ? op->k - 1 : op->k, s.fwd,
                                                                 s.u_out, s.v_in));
                }
                for (int w = 0; w < 4; ++w) {
                    DenseSweep s = sweep_desc(op, w);
                    check_mfma_bound(s, op->nl);
                    {
                        std::lock_guard<std::mutex> cl(ctx->capture_mtx);
                        build_mfma_table(op->mt[w], op->n, s.r_out, s.k_in, op->p, src[w], !s.fwd);
                    }
                    A.ptrs.push_back(op->mt[w].d);
                    A.bytes += (uint64_t)op->n * op->nl * op->mt[w].rows_pad * op->mt[w].k_pad +
                               mfma_zero_bytes(op->nl, op->mt[w].rows_pad, op->mt[w].k_pad);
                }
                op->d_lp_src = A.alloc<int8_t>(limb_plane_bytes(op->u_src, op->M, b, op->nl));
                op->d_lp_mid = A.alloc<int8_t>(limb_plane_bytes(op->u_tgt, op->Mp, b, op->nl));
            } else {
                for (int w = 0; w < 4; ++w) {
                    DenseSweep s = sweep_desc(op, w);
                    const uint32_t k4 = (s.k_in + 3) / 4 * 4;
                    op->k4[w] = k4;
                    std::vector<uint32_t> pos((uint64_t)op->n * s.r_out * k4, 0), neg(pos.size(), 0);
                    for (uint32_t i = 0; i < op->n; ++i)
                        for (uint32_t r = 0; r < s.r_out; ++r)
                            for (uint32_t kk = 0; kk < s.k_in; ++kk) {
                                // table [i][beta][a]; forward rows are beta, transpose rows a
                                uint32_t v = s.fwd ? src[w][((uint64_t)i * op->Mp + r) * op->M + kk]
                                                   : src[w][((uint64_t)i * op->Mp + kk) * op->M + r];
                                pos[((uint64_t)i * s.r_out + r) * k4 + kk] = v;
                                neg[((uint64_t)i * s.r_out + r) * k4 + kk] = v ? op->p - v : 0;
                            }
                    op->d_tab[w][0] = A.upload(pos);
                    op->d_tab[w][1] = A.upload(neg);
                }
                op->d_scaled = A.alloc<uint32_t>(op->cols * b);
            }
            op->d_mid = A.alloc<uint32_t>(op->mid * b);
        } else if (desc->kind == PGW_KIND_CYCLIC_ELIMINATED || is_deformation(desc->kind)) {
            require(desc->method != PGW_METHOD_MFMA,
                    "the cyclic operator's character blocks run on the plain path", PGW_ERR_UNSUPPORTED);
            op->method = PGW_METHOD_PLAIN;
            build_cyclic_layout(layout, *desc);
            op->rows = layout.rows;
            op->cols = layout.cols;
            if (is_deformation(desc->kind)) {
                uint32_t d0 = 0, d1 = 0;
                op->cols = deformation_cols(desc, layout.cols, d0, d1);
                const uint64_t cnt = op->rows * desc->dense_cols;
                uint32_t dlo = 0, dhi = 0;
                min_max_u32(desc->dense_block, cnt, dlo, dhi);
                require(cnt == 0 || dhi < op->p, "dense block entries must be reduced modulo p");
            }
            const uint32_t g = desc->genus, nv = g - 3;
            for (uint64_t q = 0; q < (uint64_t)nv * nv * g; ++q)
                require(desc->mu[q] < op->p, "mu.values entry is not reduced modulo the instance modulus");
            cyclic = true;
        } else {
            throw Error(PGW_ERR_INVALID, "unknown operator kind");
        }
        // preconditioners and the row-mixing plan
        validate_preconditioners(h_row_precond, op->rows, h_col_precond, op->cols, op->p);
        op->d_dr = A.upload(h_row_precond, op->rows);
        op->d_dc = A.upload(h_col_precond, op->cols);
        std::vector<uint32_t> negdc(op->cols);
        for (uint64_t q = 0; q < op->cols; ++q) negdc[q] = op->p - h_col_precond[q];
        op->d_negdc = A.upload(negdc);
        op->rounds = rowmix_rounds;
        op->pair_count = op->rows / 2;
        if (rowmix_rounds > 0 && op->pair_count > 0) {
            const uint64_t total = (uint64_t)rowmix_rounds * op->pair_count;
            op->unpaired = validate_rowmix_plan(op->rows, rowmix_rounds, h_pairs, op->p);
            // the pairs of a round are disjoint, so they may run in any order:
            // sorted by first row, half of every round's row accesses stream
            std::vector<pgw_rowmix_pair> sorted(h_pairs, h_pairs + total);
            for (uint32_t r = 0; r < rowmix_rounds; ++r)
                std::sort(sorted.begin() + (uint64_t)r * op->pair_count,
                          sorted.begin() + (uint64_t)(r + 1) * op->pair_count,
                          [](const pgw_rowmix_pair& x, const pgw_rowmix_pair& y) { return x.first < y.first; });
            op->d_pairs = A.upload(sorted);
        } else {
            op->rounds = 0;
        }
        if (cyclic) {
            CyclicMaps maps;
            build_cyclic_maps(maps, layout, *desc, op->cols, h_col_precond, h_row_precond, op->rounds,
                              h_pairs);
            layout = CyclicLayout();
            op->y_rows = maps.m1.rows;
            op->mid = op->y_rows;
            op->m1 = upload_sell(A, maps.m1, op->p);
            op->m1t = upload_sell(A, csr_transpose(maps.m1), op->p);
            maps.m1 = CsrHost();
            op->m2 = upload_sell(A, maps.m2, op->p);
            op->m2t = upload_sell(A, csr_transpose(maps.m2), op->p);
            maps.m2 = CsrHost();
            // M3 = P^T Dr P gathers each row's partners through the pairing,
            // at random in natural order (0.93 cache-line requests per entry
            // at g24) but from a few positions away in the row-mixing order
            // (0.26); PGW_CYCLIC_ORDER=0 keeps the natural order
            const char* knob = std::getenv("PGW_CYCLIC_ORDER");
            if (op->rounds > 0 && !(knob && knob[0] == '0')) {
                const std::vector<uint32_t> order = mix_order(op->rows, op->rounds, h_pairs);
                std::vector<uint32_t> pos(op->rows);
                for (uint64_t i = 0; i < op->rows; ++i) pos[order[i]] = (uint32_t)i;
                op->m3 = upload_sell(A, permute_square(maps.m3, order), op->p);
                op->d_t_row = A.upload(order);
                op->d_t_pos = A.upload(pos);
            } else {
                op->m3 = upload_sell(A, maps.m3, op->p);
            }
            op->d_y = A.alloc<uint32_t>(op->y_rows * b);
            op->d_t2 = A.alloc<uint32_t>(op->rows * b);
        }
        op->d_target = A.alloc<uint32_t>(op->rows * b);
        op->d_gram = A.alloc<uint64_t>(gram_scratch_bytes(b) / sizeof(uint64_t));
        ctx->live.fetch_add(1);
        *out = op;
    });
    if (rc != PGW_OK && op) {
        op->arena.release();
        delete op;
    }
    return rc;
}

int pgw_operator_destroy(pgw_op* op) {
    return guarded([&] {
        if (!op) return;
        PGW_HIP(hipSetDevice(op->ctx->device));
        {
            std::lock_guard<std::mutex> cl(op->ctx->capture_mtx);
            PGW_HIP(hipStreamSynchronize(op->ctx->stream));
            for (auto& e : op->stage_events) {
                (void)hipEventDestroy(e.e0);
                (void)hipEventDestroy(e.e1);
            }
        }
        op->arena.release();
        op->ctx->live.fetch_sub(1);
        delete op;
    });
}

int pgw_operator_info(const pgw_op* op, pgw_op_info* info) {
    return guarded([&] {
        require(op && info, "null operator or info");
        info->rows = op->rows;
        info->cols = op->cols;
        info->mid_rows = op->mid;
        info->device_bytes = op->arena.bytes;
        info->method = op->method;
        info->block = op->b;
    });
}

int pgw_apply_square(pgw_op* op, const uint32_t* d_in, uint32_t* d_out, void* stream) {
    return guarded([&] {
        require(op && d_in && d_out, "null operator or block");
        require(d_in != d_out, "pgw_apply_square needs distinct input and output blocks");
        PGW_HIP(hipSetDevice(op->ctx->device));
        apply_square_impl(op, d_in, d_out, pick_stream(op, stream));
    });
}

int pgw_apply_forward(pgw_op* op, const uint32_t* d_in, uint32_t* d_out, void* stream) {
    return guarded([&] {
        require(op && d_in && d_out, "null operator or block");
        PGW_HIP(hipSetDevice(op->ctx->device));
        hipStream_t st = pick_stream(op, stream);
        forward_into_target(op, d_in, st);
        PGW_HIP(hipMemcpyAsync(d_out, op->d_target, op->rows * op->b * sizeof(uint32_t),
                               hipMemcpyDeviceToDevice, st));
    });
}

int pgw_gram_upper(pgw_op* op, const uint32_t* d_x, const uint32_t* d_y, uint32_t* d_out,
                   void* stream) {
    return guarded([&] {
        require(op && d_x && d_y && d_out, "null operator or block");
        PGW_HIP(hipSetDevice(op->ctx->device));
        launch_gram(d_x, d_y, op->cols, op->b, 0, d_out, nullptr, op->d_gram, op->m,
                    pick_stream(op, stream), nullptr, 0, op->gram_mfma);
    });
}

int pgw_operator_prepare_row_mixing(pgw_op* op) {
    return guarded([&] {
        require(op != nullptr, "null operator");
        const char* knob = std::getenv("PGW_MIX_CHAINS");  // =0 keeps the pair passes (a test knob)
        if (op->kind != PGW_KIND_SYNTHETIC || op->chains.n || !(op->rounds == 1 || op->rounds == 2) ||
            op->rows >= (1ULL << 31) - 1 || (knob && knob[0] == '0'))
            return;
        PGW_HIP(hipSetDevice(op->ctx->device));
        const uint64_t pc = op->pair_count, total = (uint64_t)op->rounds * pc;
        std::vector<pgw_rowmix_pair> pairs(total);
        std::vector<uint32_t> dr(op->rows);
        {
            std::lock_guard<std::mutex> cl(op->ctx->capture_mtx);
            PGW_HIP(hipStreamSynchronize(op->ctx->stream));
            if (total)
                PGW_HIP(hipMemcpy(pairs.data(), op->d_pairs, total * sizeof(pgw_rowmix_pair), hipMemcpyDeviceToHost));
            PGW_HIP(hipMemcpy(dr.data(), op->d_dr, op->rows * sizeof(uint32_t), hipMemcpyDeviceToHost));
        }
        MixChains mc = mix_chains(op->rows, op->rounds, pairs.data(), dr.data());
        pgw::MixChainsDevice c;
        c.n = mc.row.size();
        c.row = op->arena.upload(mc.row);
        c.a = op->arena.upload(mc.a);
        c.b = op->arena.upload(mc.b);
        c.dr = op->arena.upload(mc.dr);
        c.flag = op->arena.upload(mc.flag);
        op->d_target2 = op->arena.alloc<uint32_t>(op->rows * op->b);
        op->chains = c;
    });
}

int pgw_operator_set_gram(pgw_op* op, uint32_t method) {
    return guarded([&] {
        require(op != nullptr, "null operator");
        if (method == PGW_METHOD_AUTO) {
            op->gram_mfma = op->method == PGW_METHOD_MFMA;
        } else if (method == PGW_METHOD_MFMA) {
            require(op->b == 4 || op->b == 8 || op->b == 16, "MFMA Gram projection needs b in {4, 8, 16}",
                    PGW_ERR_UNSUPPORTED);
            op->gram_mfma = true;
        } else {
            require(method == PGW_METHOD_PLAIN, "unknown Gram method");
            op->gram_mfma = false;
        }
    });
}

int pgw_gram_upper_noalloc(pgw_ctx* ctx, uint32_t prime, uint32_t block, uint64_t rows,
                           const uint32_t* d_x, const uint32_t* d_y, uint32_t* d_out, void* stream) {
    return guarded([&] {
        require(ctx && d_x && d_y && d_out, "null context or block");
        require(prime > 2 && prime < (1u << 31), "prime must satisfy 2 < p < 2^31");
        require(block >= 1 && block <= 16, "block width must be in 1..16");
        require(rows < (1ULL << 32), "Gram projection supports fewer than 2^32 rows");
        PGW_HIP(hipSetDevice(ctx->device));
        launch_gram_noalloc(ctx->device, d_x, d_y, rows, block, d_out, make_modp(prime),
                            stream ? (hipStream_t)stream : ctx->stream);
    });
}

int pgw_sequence_run(pgw_op* op, const uint32_t* h_v, uint64_t max_len, uint32_t* h_seq,
                     uint32_t* h_curv, uint64_t* out_len, pgw_progress_fn progress, void* user) {
    uint32_t* d_B = nullptr;
    uint32_t* d_D = nullptr;
    uint32_t* d_seq = nullptr;
    uint32_t* d_pos = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int rc = guarded([&] {
        require(op && h_v && h_seq && h_curv, "null operator or buffer");
        require(max_len >= 1, "sequence length must be positive");
        PGW_HIP(hipSetDevice(op->ctx->device));
        hipStream_t st = op->ctx->stream;
        const uint64_t N = op->cols, b = op->b, ut = b * (b + 1) / 2;
        const uint64_t rounds = (max_len - 1 + 3) / 4;
        const uint64_t L = 1 + 4 * rounds;
        {
            uint32_t lo, hi;
            min_max_u32(h_v, N * b, lo, hi);
            require(N * b == 0 || hi < op->p, "V entries must be reduced modulo p");
        }
        {
            std::lock_guard<std::mutex> cl(op->ctx->capture_mtx);
            PGW_HIP(hipMalloc(&d_B, N * b * sizeof(uint32_t)));
            PGW_HIP(hipMalloc(&d_D, N * b * sizeof(uint32_t)));
            PGW_HIP(hipMalloc(&d_seq, L * ut * sizeof(uint32_t)));
            PGW_HIP(hipMalloc(&d_pos, sizeof(uint32_t)));
            // the next sequence position, read on the device; a fused round
            // moves it on at its start, so it starts one round back
            unsigned nb_unused = 0;
            const uint32_t one = fused_round_plan(op, &nb_unused) ? 1u - 4u : 1u;
            PGW_HIP(hipMemcpy(d_pos, &one, sizeof(uint32_t), hipMemcpyHostToDevice));
        }
        PGW_HIP(hipMemcpyAsync(d_B, h_v, N * b * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        launch_gram(d_B, d_B, N, op->b, 0, d_seq, nullptr, op->d_gram, op->m, st, nullptr, 0, op->gram_mfma);  // V^T V
        PGW_HIP(hipStreamSynchronize(st));
        // a long loop of a synthetic operator builds its row-mixing chains
        // first (host work linear in the rows, repaid many times over)
        if (rounds >= kPrepareRounds && op->kind == PGW_KIND_SYNTHETIC && !op->chains.n) {
            // on failure (device memory for the chains) the loop keeps the pair passes
            (void)pgw_operator_prepare_row_mixing(op);
        }
        if (rounds) {
            // one round (cuprym_cyclic.cu:2439-2449) captured once and replayed:
            // D = S B; B^T D, D^T D; B = S D; B^T D, B^T B; pos += 4
            std::lock_guard<std::mutex> cl(op->ctx->capture_mtx);
            PGW_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                unsigned nb = 0;
                if (fused_round_plan(op, &nb)) {
                    uint64_t* part_a = op->d_gram;
                    uint64_t* part_b = op->d_gram + gram_scratch_bytes(op->b) / 2 / sizeof(uint64_t);
                    const size_t half = gram_scratch_bytes(op->b) / 2;
                    unsigned na = 0, nb2 = 0;
                    cyclic_front(op, d_B, st, d_pos);
                    require(launch_sell_spmm_gram(op->m1t, op->d_y, d_D, d_B, 1, d_seq, d_seq + ut, part_a, half,
                                                  d_pos, 0, op->m, st, &na),
                            "fused projection declined a planned round");
                    cyclic_front(op, d_D, st);
                    require(launch_sell_spmm_gram(op->m1t, op->d_y, d_B, d_D, 2, d_seq, d_seq + ut, part_b, half,
                                                  d_pos, 2, op->m, st, &nb2),
                            "fused projection declined a planned round");
                    launch_gram_finish_round(part_a, na, part_b, nb2, d_seq, d_seq + ut, op->m, d_pos, st);
                } else {
                    apply_and_project(op, d_B, d_D, 1, d_seq, d_seq + ut, d_pos, 0, st);
                    apply_and_project(op, d_D, d_B, 2, d_seq, d_seq + ut, d_pos, 2, st);
                    launch_advance_pos(d_pos, 4, st);
                }
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            PGW_HIP(hipStreamEndCapture(st, &graph));
            PGW_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        }
        // rows [0, copied) of h_seq hold their final values
        uint64_t copied = 0;
        for (uint64_t r = 0; r < rounds; ++r) {
            PGW_HIP(hipGraphLaunch(exec, st));
            if ((r + 1) % 64 == 0 || r + 1 == rounds) {
                PGW_HIP(hipStreamSynchronize(st));  // also bounds the launch queue
                if (progress) {
                    const uint64_t done = 1 + 4 * (r + 1);
                    PGW_HIP(hipMemcpy(h_seq + copied * ut, d_seq + copied * ut, (done - copied) * ut * sizeof(uint32_t),
                                      hipMemcpyDeviceToHost));
                    copied = done;
                    progress(done, L, user);
                }
            }
        }
        if (copied < L)
            PGW_HIP(hipMemcpyAsync(h_seq + copied * ut, d_seq + copied * ut, (L - copied) * ut * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, st));
        PGW_HIP(hipMemcpyAsync(h_curv, d_B, N * b * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        PGW_HIP(hipStreamSynchronize(st));
        if (out_len) *out_len = L;
    });
    // a failed call may leave work queued that still reads these buffers
    if (op) (void)hipStreamSynchronize(op->ctx->stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (op) {
        std::lock_guard<std::mutex> cl(op->ctx->capture_mtx);
        for (uint32_t* p : {d_B, d_D, d_seq, d_pos})
            if (p) (void)hipFree(p);
    }
    return rc;
}

int pgw_kernel_vectors(pgw_op* op, const uint32_t* h_v, const uint32_t* h_gens, uint64_t ngen,
                       uint32_t* h_d, uint32_t* h_product) {
    uint32_t* d_B = nullptr;
    uint32_t* d_N = nullptr;
    uint32_t* d_D = nullptr;
    uint32_t* d_G = nullptr;
    int rc = guarded([&] {
        require(op && h_v && h_gens && h_d && h_product, "null operator or buffer");
        require(ngen >= 1, "generator file does not contain any generator matrices");
        PGW_HIP(hipSetDevice(op->ctx->device));
        hipStream_t st = op->ctx->stream;
        const uint64_t N = op->cols, b = op->b;
        PGW_HIP(hipMalloc(&d_B, N * b * sizeof(uint32_t)));
        PGW_HIP(hipMalloc(&d_N, N * b * sizeof(uint32_t)));
        PGW_HIP(hipMalloc(&d_D, N * b * sizeof(uint32_t)));
        PGW_HIP(hipMalloc(&d_G, ngen * b * b * sizeof(uint32_t)));
        PGW_HIP(hipMemcpyAsync(d_B, h_v, N * b * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        PGW_HIP(hipMemcpyAsync(d_G, h_gens, ngen * b * b * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        PGW_HIP(hipMemsetAsync(d_D, 0, N * b * sizeof(uint32_t), st));
        for (uint64_t kk = 0; kk < ngen; ++kk) {
            launch_accumulate_generator(d_B, d_G + kk * b * b, d_D, N, op->b, op->m, st);
            if (kk + 1 < ngen) {
                apply_square_impl(op, d_B, d_N, st);
                std::swap(d_B, d_N);
            }
            if ((kk + 1) % 256 == 0) PGW_HIP(hipStreamSynchronize(st));
        }
        forward_into_target(op, d_D, st);
        PGW_HIP(hipMemcpyAsync(h_d, d_D, N * b * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        PGW_HIP(hipMemcpyAsync(h_product, op->d_target, op->rows * b * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, st));
        PGW_HIP(hipStreamSynchronize(st));
    });
    for (uint32_t* p : {d_B, d_N, d_D, d_G})
        if (p) (void)hipFree(p);
    return rc;
}

int pgw_operator_autotune(pgw_op* op, uint32_t reps, uint32_t* h_choice, double* h_ms) {
    return guarded([&] {
        require(op != nullptr, "null operator");
        autotune_sweeps(op, reps, 0, mfma_config_count() - 1, h_choice, h_ms);
    });
}

namespace {

// one profiled apply: per stage, summed milliseconds, bytes and launches
void profile_one_apply(pgw_op* op, const uint32_t* in, uint32_t* out, hipStream_t st, double* ms,
                       double* bytes, uint32_t* launches) {
    for (auto& e : op->stage_events) {
        (void)hipEventDestroy(e.e0);
        (void)hipEventDestroy(e.e1);
    }
    op->stage_events.clear();
    op->profile = true;
    try {
        apply_square_impl(op, in, out, st);
    } catch (...) {
        op->profile = false;
        throw;
    }
    op->profile = false;
    PGW_HIP(hipStreamSynchronize(st));
    for (int q = 0; q < PGW_STAGES; ++q) {
        ms[q] = 0;
        bytes[q] = 0;
        launches[q] = 0;
    }
    for (auto& e : op->stage_events) {
        float t = 0;
        PGW_HIP(hipEventElapsedTime(&t, e.e0, e.e1));
        ms[e.stage] += t;
        bytes[e.stage] += e.bytes;
        launches[e.stage] += 1;
        (void)hipEventDestroy(e.e0);
        (void)hipEventDestroy(e.e1);
    }
    op->stage_events.clear();
}

}  // namespace

int pgw_bench_apply(pgw_op* op, uint32_t* d_a, uint32_t* d_b, uint32_t iters, double* total_ms,
                    double* sweep_avg_ms, uint32_t* sweeps_per_apply) {
    return guarded([&] {
        require(op && d_a && d_b && iters >= 1, "null operator/block or zero iterations");
        PGW_HIP(hipSetDevice(op->ctx->device));
        hipStream_t st = op->ctx->stream;
        hipEvent_t t0, t1;
        PGW_HIP(hipEventCreate(&t0));
        PGW_HIP(hipEventCreate(&t1));
        PGW_HIP(hipStreamSynchronize(st));
        // timed region: iters block-applies, no per-kernel events
        PGW_HIP(hipEventRecord(t0, st));
        for (uint32_t it = 0; it < iters; ++it)
            apply_square_impl(op, (it % 2) ? d_b : d_a, (it % 2) ? d_a : d_b, st);
        PGW_HIP(hipEventRecord(t1, st));
        PGW_HIP(hipEventSynchronize(t1));
        float ms = 0;
        PGW_HIP(hipEventElapsedTime(&ms, t0, t1));
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        if (total_ms) *total_ms = ms;
        // per-launch sweep durations: one more apply with events around each stage
        double sms[PGW_STAGES], sbytes[PGW_STAGES];
        uint32_t sl[PGW_STAGES];
        profile_one_apply(op, (iters % 2) ? d_b : d_a, (iters % 2) ? d_a : d_b, st, sms, sbytes, sl);
        if (sweep_avg_ms) *sweep_avg_ms = sl[0] ? sms[0] / sl[0] : 0.0;
        if (sweeps_per_apply) *sweeps_per_apply = sl[0];
    });
}

int pgw_bench_stages(pgw_op* op, uint32_t* d_a, uint32_t* d_b, uint32_t iters, double* h_ms,
                     double* h_bytes, uint32_t* h_launches) {
    return guarded([&] {
        require(op && d_a && d_b && iters >= 1 && h_ms && h_bytes && h_launches,
                "null operator, block or output, or zero iterations");
        PGW_HIP(hipSetDevice(op->ctx->device));
        hipStream_t st = op->ctx->stream;
        for (int q = 0; q < PGW_STAGES; ++q) h_ms[q] = h_bytes[q] = 0, h_launches[q] = 0;
        double ms[PGW_STAGES], bytes[PGW_STAGES];
        uint32_t launches[PGW_STAGES];
        for (uint32_t it = 0; it < iters; ++it) {
            profile_one_apply(op, (it % 2) ? d_b : d_a, (it % 2) ? d_a : d_b, st, ms, bytes, launches);
            for (int q = 0; q < PGW_STAGES; ++q) {
                h_ms[q] += ms[q] / iters;
                h_bytes[q] = bytes[q];
                h_launches[q] = launches[q];
            }
        }
    });
}

}  // extern "C"

// pgw_internal.h -- shared declarations of libpgw (host structs, device helpers).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/pgw.h"

namespace pgw {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

void set_last_error(const std::string& msg);

#define PGW_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t pgw_e_ = (call);                                                     \
        if (pgw_e_ != hipSuccess)                                                       \
            throw ::pgw::Error(PGW_ERR_HIP, std::string(#call " failed: ") +           \
                                                hipGetErrorString(pgw_e_));             \
    } while (0)

#define PGW_CHECK_LAUNCH(name)                                                          \
    do {                                                                                \
        hipError_t pgw_e_ = hipGetLastError();                                          \
        if (pgw_e_ != hipSuccess)                                                       \
            throw ::pgw::Error(PGW_ERR_HIP, std::string(name " launch failed: ") +     \
                                                hipGetErrorString(pgw_e_));             \
    } while (0)

inline void require(bool cond, const std::string& msg, int status = PGW_ERR_INVALID) {
    if (!cond) throw Error(status, msg);
}
// literal messages: no string is built unless the check fails (hot loops)
inline void require(bool cond, const char* msg, int status = PGW_ERR_INVALID) {
    if (!cond) throw Error(status, msg);
}

// smallest and largest entry of x[0..n) (one vectorisable pass); {~0u, 0} for n = 0
inline void min_max_u32(const uint32_t* x, uint64_t n, uint32_t& lo, uint32_t& hi) {
    uint32_t a = ~0u, b = 0;
    for (uint64_t q = 0; q < n; ++q) {
        a = x[q] < a ? x[q] : a;
        b = x[q] > b ? x[q] : b;
    }
    lo = a;
    hi = b;
}

// ---------------------------------------------------------------------------
// modular arithmetic (p prime, 2 < p < 2^31)
// ---------------------------------------------------------------------------

struct ModP {
    uint32_t p;
    uint32_t r32;     // 2^32 mod p
    uint64_t barrett; // floor(2^64 / p)
};

inline ModP make_modp(uint32_t p) {
    ModP m;
    m.p = p;
    m.r32 = (uint32_t)((1ULL << 32) % p);
    m.barrett = (uint64_t)(((unsigned __int128)1 << 64) / p);
    return m;
}

// x mod p for any 64-bit x: Barrett with m = floor(2^64/p) leaves r < 3p.
__host__ __device__ inline uint32_t mod_u64(uint64_t x, const ModP& m) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t q = __umul64hi(x, m.barrett);
#else
    uint64_t q = (uint64_t)(((unsigned __int128)x * m.barrett) >> 64);
#endif
    uint64_t r = x - q * (uint64_t)m.p;
    if (r >= m.p) r -= m.p;
    if (r >= m.p) r -= m.p;
    return (uint32_t)r;
}

__host__ __device__ inline uint32_t mulmod(uint32_t a, uint32_t b, const ModP& m) {
    return mod_u64((uint64_t)a * b, m);
}

__host__ __device__ inline uint32_t addmod(uint32_t a, uint32_t b, uint32_t p) {
    uint32_t s = a + b;  // a, b < p < 2^31: no wrap
    return s >= p ? s - p : s;
}

// (hi * 2^32 + lo) mod p for the split accumulators of the plain kernels.
__host__ __device__ inline uint32_t mod_split(uint64_t hi, uint64_t lo, const ModP& m) {
    uint64_t h = mod_u64(hi, m);
    return mod_u64(h * m.r32 + mod_u64(lo, m), m);
}

// ---------------------------------------------------------------------------
// exterior-algebra basis: lexicographic k-subsets of [0,n) (n < 64)
// ---------------------------------------------------------------------------

// Uploads the binomial table C(a,b), a,b < 64, to the device constant memory
// of the kernel module (once per device).
void upload_binomials(int device);
uint64_t host_binom(uint32_t n, uint32_t k);

template <typename Binom>
__host__ __device__ inline uint64_t unrank_mask(uint64_t rank, int n, int k, const Binom& C) {
    // greedy lexicographic unrank (same order as the reference's unrank_subset)
    uint64_t mask = 0;
    int start = 0;
    for (int pos = 0; pos < k; ++pos) {
        int remaining = k - pos - 1;
        for (int value = start; value < n; ++value) {
            uint64_t count = C(n - value - 1, remaining);
            if (rank < count) {
                mask |= 1ULL << value;
                start = value + 1;
                break;
            }
            rank -= count;
        }
    }
    return mask;
}

// Lexicographic rank of a k-subset given as a bit mask:
// rank = sum_pos [ C(n - s_{pos-1} - 1, k - pos) - C(n - s_pos, k - pos) ]
// (hockey-stick sum of the greedy unrank counts, s_{-1} = -1).
template <typename Binom>
__host__ __device__ inline uint64_t rank_mask(uint64_t mask, int n, int k, const Binom& C) {
    uint64_t r = 0;
    int prev = -1, pos = 0;
    while (mask) {
        int s = __builtin_ctzll(mask);
        mask &= mask - 1;
        r += C(n - prev - 1, k - pos) - C(n - s, k - pos);
        prev = s;
        ++pos;
    }
    return r;
}

// C(a, b) for 0 <= b <= a < 64 (Pascal's triangle, built once; every entry
// fits in 64 bits), row-major 64 x 64
const uint64_t* host_binom_table();

struct HostBinom {
    const uint64_t* t = host_binom_table();
    uint64_t operator()(int a, int b) const {
        if (a < 0 || b < 0 || b > a) return 0ULL;
        return a < 64 ? t[a * 64 + b] : host_binom((uint32_t)a, (uint32_t)b);
    }
};

inline unsigned host_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return std::max(1u, std::min(16u, h ? h : 1u));
}

// f(lo, hi) over [0, n) split across host threads
template <typename F>
void parallel_rows(uint64_t n, F&& f) {
    const unsigned T = n < 4096 ? 1u : host_threads();
    if (T == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    const uint64_t step = (n + T - 1) / T;
    for (unsigned q = 0; q < T; ++q) {
        const uint64_t lo = q * step, hi = std::min(n, lo + step);
        if (lo < hi) th.emplace_back([&f, lo, hi] { f(lo, hi); });
    }
    for (auto& t : th) t.join();
}

}  // namespace pgw

// ---------------------------------------------------------------------------
// host-visible objects
// ---------------------------------------------------------------------------

struct pgw_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t* d_binom = nullptr;  // C(a,b), a,b < 64, row-major
    // Held while a batch captures a graph and around the allocation, upload
    // and free calls that may run on other threads meanwhile (operator
    // create/destroy, batch add/result/cancel): such calls during a capture
    // invalidate it.
    std::mutex capture_mtx;
    // operators and batches created on this context and not yet destroyed:
    // the context is not freed while any of them could still use it
    std::atomic<uint64_t> live{0};
};

namespace pgw {

// Host-side layout of one differential (built like build_common_incidences_generic
// + build_sector_layout_generic, cuda/cuprym_cyclic.cu:631-790).
struct PhiHost {
    uint32_t nI = 0, nB = 0, nA = 0, r = 1;
    std::vector<uint32_t> mu;
    std::vector<uint16_t> wV, wA0, wA1;
    std::vector<uint32_t> a0_off, a1_off;
    std::vector<uint16_t> a0_idx, a1_idx;
    uint64_t row_subsets = 0, col_subsets = 0;
    uint32_t row_width = 0, col_width = 0;
    std::vector<uint16_t> row_subset_w, col_subset_w;
    std::vector<uint32_t> row_inc_code, row_inc_rank;  // code = i | neg<<31
    std::vector<uint32_t> col_inc_code, col_inc_rank;
    std::vector<uint32_t> row_off, row_rank, col_off, col_rank;
    std::vector<uint16_t> row_beta, col_a;
};

void build_char_map(const std::vector<uint16_t>& w, uint32_t r, std::vector<uint32_t>& off,
                    std::vector<uint16_t>& idx);
void build_phi_host(PhiHost& h, uint32_t nI, uint32_t nB, uint32_t nA, uint32_t r,
                    const uint32_t* mu, const uint16_t* wV, const uint16_t* wA0,
                    const uint16_t* wA1, const std::vector<uint32_t>& original, uint32_t degree,
                    uint32_t sector);

// Cyclic eliminated layout sizes and z layout.
struct CyclicLayout {
    PhiHost dm, d2;
    std::vector<uint32_t> z_off, z_rank;
    std::vector<uint16_t> z_kappa;
    uint64_t x_cols = 0, z_count = 0, rows = 0, cols = 0;
};
void build_cyclic_layout(CyclicLayout& L, const pgw_operator_desc& d);

// Entries (column, coefficient) of output row `row` of a character-restricted
// differential's forward gather (the index rules of
// cyclic_phi_forward_scaled_kernel, cuda/cyclic_cuda_helpers.h:145-204): for
// each incidence (i, s' + i) and each a in the character block of
// w(beta) - w(i), the coefficient +-mu[i][beta][a] on input column
// colbase(s' + i) + local_a.
template <typename Emit>
void phi_forward_row(const PhiHost& h, uint64_t row, uint32_t p, Emit&& emit) {
    const uint32_t beta = h.row_beta[row];
    const uint32_t bw = h.wA1[beta];
    const uint64_t s = h.row_rank[row];
    for (uint32_t pos = 0; pos < h.row_width; ++pos) {
        const uint32_t code = h.row_inc_code[s * h.row_width + pos];
        const uint32_t i = code & 0x7fffffffu;
        const bool neg = code >> 31;
        const uint32_t q = (bw + h.r - (h.wV[i] % h.r)) % h.r;
        const uint32_t start = h.a0_off[q], end = h.a0_off[q + 1];
        const uint64_t base = h.col_off[h.row_inc_rank[s * h.row_width + pos]];
        const uint32_t* murow = h.mu.data() + ((uint64_t)i * h.nB + beta) * h.nA;
        for (uint32_t idx = start; idx < end; ++idx) {
            const uint32_t t = murow[h.a0_idx[idx]];
            if (t == 0) continue;
            emit(base + (idx - start), neg ? p - t : t);
        }
    }
}

// Descriptor and input checks shared by the device operator and the host
// apply (pgw_api.cpp).
void validate_desc_common(const pgw_operator_desc* d);
bool is_deformation(uint32_t kind);
// source columns of a deformation variant over a base of `base` columns;
// drop0 < drop1 are the replacement's dropped base columns
uint64_t deformation_cols(const pgw_operator_desc* d, uint64_t base, uint32_t& drop0, uint32_t& drop1);
// disjointness, range and coefficients of a row-mixing plan; returns the row
// left out of the rounds when rows is odd, else -1
int64_t validate_rowmix_plan(uint64_t rows, uint32_t rounds, const pgw_rowmix_pair* pairs, uint32_t p);
void validate_preconditioners(const uint32_t* dr, uint64_t rows, const uint32_t* dc, uint64_t cols,
                              uint32_t p);

// Explicit sparse maps (pgw_sparse.cpp): CSR with sorted columns and nonzero
// canonical coefficients, and its sliced-ELL form for the device (slices of
// kSellHeight rows; entry k of row r of slice s at ent[2 (slice[s] +
// k kSellHeight + r % kSellHeight)] = column, coefficient; rows shorter than
// their slice's widest are padded with (0, 0)).
constexpr uint64_t kSellHeight = 64;
struct CsrHost {
    uint64_t rows = 0, cols = 0;
    std::vector<uint64_t> ptr;
    std::vector<uint32_t> col, val;
};
struct SellHost {
    uint64_t rows = 0, cols = 0, nnz = 0;
    std::vector<uint64_t> slice;
    std::vector<uint32_t> ent;
};
CsrHost csr_transpose(const CsrHost& A);
SellHost csr_to_sell(const CsrHost& A);
// The eliminated cyclic operator (or a deformation variant) with Dc, the
// elimination and P^T Dr P folded in: S = M1^T M2^T M3 M2 M1 (pgw_sparse.cpp).
struct CyclicMaps {
    CsrHost m1, m2, m3;
};
void build_cyclic_maps(CyclicMaps& out, const CyclicLayout& L, const pgw_operator_desc& d,
                       uint64_t cols, const uint32_t* dc, const uint32_t* dr, uint32_t rounds,
                       const pgw_rowmix_pair* pairs);
// An order of the `rows` target rows in which the row-mixing rounds are
// local: every pair of every round links two rows, and the order walks the
// linked components (with two rounds: paths and cycles, so that each row's
// partners, and theirs, sit a few positions away).  order[i] = the row at
// position i.
std::vector<uint32_t> mix_order(uint64_t rows, uint32_t rounds, const pgw_rowmix_pair* pairs);
// The target rows as the chains the row-mixing rounds link (rounds <= 2:
// paths and cycles), for the one-pass P^T Dr P (k_mix_chain): position i holds
// target row row[i]; flag[i] bits 0-1 give the round of the pair linking
// positions i and i + 1 (kMixNone: none), bit 2 says row[i] is that pair's
// `first` (else row[i + 1] is), bit 3 marks a ghost (a copy of a cycle row
// placed so that every row's neighbours within kMixHalo positions are at
// hand; it is read, never written); a[i], b[i] are that pair's coefficients
// and dr[i] = Dr[row[i]].  Every target row appears exactly once unghosted.
constexpr uint32_t kMixNone = 3, kMixFirst = 4, kMixGhost = 8, kMixHalo = 4;
struct MixChains {
    std::vector<uint32_t> row, a, b, dr;
    std::vector<uint8_t> flag;
};
MixChains mix_chains(uint64_t rows, uint32_t rounds, const pgw_rowmix_pair* pairs, const uint32_t* dr);
// A with rows and columns renumbered: row i of the result is row order[i] of
// A, column c becomes the position of c in `order` (square A).
CsrHost permute_square(const CsrHost& A, const std::vector<uint32_t>& order);
// out = S in on a cols x w host block, matrix-free from the layout (pgw_cpu.cpp)
void cyclic_apply_host(const CyclicLayout& L, const pgw_operator_desc& d, uint64_t cols, const uint32_t* dr,
                       const uint32_t* dc, uint32_t rounds, const pgw_rowmix_pair* pairs, const uint32_t* in,
                       uint32_t w, uint32_t* out);

}  // namespace pgw

// pgw_kernels.hip -- gfx950 kernels of the block Wiedemann sequence generator.
//
// Two implementations of the Koszul sweep (the matrix-free gather of the
// reference's cyclic_phi_forward_scaled_kernel / cyclic_phi_transpose_kernel,
// cuda/cyclic_cuda_helpers.h:145-266):
//   * plain: u64 modular multiply-accumulate on the vector ALUs, b lanes of a
//     wavefront per output row so that each gathered input row (b words) is one
//     coalesced segment;
//   * mfma: the dense multiplication-table contraction T_i . X_{s} as int8 MFMA
//     (v_mfma_i32_16x16x64_i8) on balanced base-256 digits, int32 accumulation
//     per digit shift, one exact reduction mod p in the epilogue.
// Every kernel is exact: results are canonical residues, identical whatever the
// summation order, because no accumulator can overflow (bounds stated inline).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <utility>

#include "pgw_kernels.h"

namespace pgw {

__constant__ uint64_t c_binom[64][64];

struct DeviceBinom {
    __device__ uint64_t operator()(int a, int b) const {
        return (a < 0 || b < 0 || b > a) ? 0ULL : c_binom[a][b];
    }
};

void upload_binomials(int device) {
    static std::mutex mtx;
    static std::vector<int> done;
    std::lock_guard<std::mutex> lock(mtx);
    if (std::find(done.begin(), done.end(), device) != done.end()) return;
    static uint64_t table[64][64];
    for (int a = 0; a < 64; ++a)
        for (int b = 0; b < 64; ++b) {
            if (b > a) { table[a][b] = 0; continue; }
            int kk = std::min(b, a - b);
            unsigned __int128 out = 1;
            for (int i = 0; i < kk; ++i) out = (out * (unsigned)(a - i)) / (unsigned)(i + 1);
            // entries above 2^64 are never used (subset counts are checked to fit u32)
            table[a][b] = out > (unsigned __int128)UINT64_MAX ? UINT64_MAX : (uint64_t)out;
        }
    PGW_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_binom), table, sizeof(table), 0, hipMemcpyHostToDevice));
    done.push_back(device);
}

static inline unsigned grid_for(uint64_t threads, unsigned block = 256) {
    uint64_t g = (threads + block - 1) / block;
    return (unsigned)std::max<uint64_t>(1, g);
}

// ===========================================================================
// plain dense-table sweep (synthetic operator, r = 1)
// ===========================================================================
//
// out[(u*R_out + r)*B + c] = scale[u*R_out+r] * sum_{i ~ u} sum_k tab_i^{+/-}[r][k] in[(nbr*K_in + k)*B + c]
// FWD: u is a (k-1)-subset, i ranges over i not in u, nbr = rank(u + {i}) among k-subsets.
// TRN: u is a k-subset, i ranges over i in u, nbr = rank(u - {i}) among (k-1)-subsets.
// The sign (-1)^{#{x in u : x < i}} selects the negated table copy.
// Accumulation: BIGP (p < 2^31) groups 4 products (< 4(p-1)^2 < 2^64) and splits
// them into 64-bit low/high word sums; SMALLP (p < 2^16) sums 32-bit products.

template <int B, bool FWD, bool BIGP>
__global__ __launch_bounds__(256) void k_dense_plain(
    const uint32_t* __restrict__ tab, const uint32_t* __restrict__ tab_neg, uint32_t n,
    uint32_t deg_out, uint64_t u_out, uint32_t r_out, uint32_t k_in, uint32_t k4,
    const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
    const uint32_t* __restrict__ scale, ModP m) {
    __shared__ uint32_t s_code[64];
    __shared__ uint32_t s_nbr[64];
    __shared__ int s_cnt;
    const uint64_t u = blockIdx.x;
    if (u >= u_out) return;
    if (threadIdx.x == 0) {
        DeviceBinom C;
        uint64_t mask = unrank_mask(u, (int)n, (int)deg_out, C);
        int cnt = 0;
        for (uint32_t i = 0; i < n; ++i) {
            bool bit = (mask >> i) & 1ULL;
            if (FWD ? bit : !bit) continue;
            uint32_t neg = (uint32_t)__popcll(mask & ((1ULL << i) - 1)) & 1u;
            uint64_t nb = FWD ? rank_mask(mask | (1ULL << i), (int)n, (int)deg_out + 1, C)
                              : rank_mask(mask & ~(1ULL << i), (int)n, (int)deg_out - 1, C);
            s_code[cnt] = i | (neg << 31);
            s_nbr[cnt] = (uint32_t)nb;
            ++cnt;
        }
        s_cnt = cnt;
    }
    __syncthreads();
    constexpr int RPB = 256 / B;
    const int c = threadIdx.x % B;
    const int rl = threadIdx.x / B;
    if (rl >= RPB) return;
    const int cnt = s_cnt;
    const uint32_t k_main = k_in & ~3u;
    for (uint32_t r = rl; r < r_out; r += RPB) {
        uint64_t lo = 0, hi = 0;
        for (int w = 0; w < cnt; ++w) {
            const uint32_t code = s_code[w];
            const uint32_t i = code & 0x7fffffffu;
            const uint32_t* trow = ((code >> 31) ? tab_neg : tab) + ((uint64_t)i * r_out + r) * k4;
            const uint32_t* x = in + (uint64_t)s_nbr[w] * k_in * B + c;
            uint32_t k = 0;
            for (; k < k_main; k += 4) {
                const uint4 t = *reinterpret_cast<const uint4*>(trow + k);
                const uint32_t x0 = x[(k + 0) * B], x1 = x[(k + 1) * B];
                const uint32_t x2 = x[(k + 2) * B], x3 = x[(k + 3) * B];
                if (BIGP) {
                    uint64_t s = (uint64_t)t.x * x0 + (uint64_t)t.y * x1 + (uint64_t)t.z * x2 +
                                 (uint64_t)t.w * x3;
                    lo += (uint32_t)s;
                    hi += s >> 32;
                } else {
                    lo += (uint64_t)(t.x * x0) + (uint64_t)(t.y * x1) + (uint64_t)(t.z * x2) +
                          (uint64_t)(t.w * x3);
                }
            }
            for (; k < k_in; ++k) {
                uint64_t s = (uint64_t)trow[k] * x[k * B];
                lo += (uint32_t)s;
                hi += s >> 32;
            }
        }
        uint32_t v = mod_split(hi, lo, m);
        const uint64_t row = u * r_out + r;
        if (scale) v = mulmod(v, scale[row], m);
        out[row * B + c] = v;
    }
}

template <int B>
static void dense_plain_b(const DenseSweep& s, const uint32_t* tab, const uint32_t* tab_neg,
                          uint32_t k4, const uint32_t* in, uint32_t* out, const uint32_t* scale,
                          const ModP& m, hipStream_t st) {
    const bool big = m.p >= 65536u;
    const uint32_t deg_out = s.fwd ? s.k - 1 : s.k;
    dim3 grid((unsigned)s.u_out), block(256);
#define PGW_DP(F, BIG)                                                                      \
    hipLaunchKernelGGL((k_dense_plain<B, F, BIG>), grid, block, 0, st, tab, tab_neg, s.n,  \
                       deg_out, s.u_out, s.r_out, s.k_in, k4, in, out, scale, m)
    if (s.fwd) { if (big) PGW_DP(true, true); else PGW_DP(true, false); }
    else { if (big) PGW_DP(false, true); else PGW_DP(false, false); }
#undef PGW_DP
    PGW_CHECK_LAUNCH("dense plain sweep");
}

#define PGW_SWITCH_B(b, CALL)                                                               \
    switch (b) {                                                                            \
        case 1: CALL(1); break;   case 2: CALL(2); break;   case 3: CALL(3); break;         \
        case 4: CALL(4); break;   case 5: CALL(5); break;   case 6: CALL(6); break;         \
        case 7: CALL(7); break;   case 8: CALL(8); break;   case 9: CALL(9); break;         \
        case 10: CALL(10); break; case 11: CALL(11); break; case 12: CALL(12); break;       \
        case 13: CALL(13); break; case 14: CALL(14); break; case 15: CALL(15); break;       \
        case 16: CALL(16); break;                                                           \
        default: throw Error(PGW_ERR_UNSUPPORTED, "block width must be in 1..16");          \
    }

void launch_dense_plain(const DenseSweep& s, const uint32_t* tab, const uint32_t* tab_neg,
                        uint32_t k4, const uint32_t* in, uint32_t* out, const uint32_t* scale,
                        uint32_t b, const ModP& m, hipStream_t st) {
    require(s.u_out < (1ULL << 31), "too many output subsets for one launch");
#define PGW_CALL(BB) dense_plain_b<BB>(s, tab, tab_neg, k4, in, out, scale, m, st)
    PGW_SWITCH_B(b, PGW_CALL)
#undef PGW_CALL
}

// ===========================================================================
// plain character-restricted sweeps (reference's cyclic operator)
// ===========================================================================

struct PhiArgs {
    uint32_t nB, nA, r, row_width, col_width;
    uint64_t rows, cols;
    const uint32_t* mu;
    const uint16_t* wV;
    const uint16_t* wA0;
    const uint16_t* wA1;
    const uint32_t* a0_off;
    const uint16_t* a0_idx;
    const uint32_t* a1_off;
    const uint16_t* a1_idx;
    const uint32_t* row_off;
    const uint32_t* row_rank;
    const uint16_t* row_beta;
    const uint32_t* col_off;
    const uint32_t* col_rank;
    const uint16_t* col_a;
    const uint32_t* row_inc;
    const uint32_t* row_inc_rank;
    const uint32_t* col_inc;
    const uint32_t* col_inc_rank;
};

static PhiArgs phi_args(const PhiDevice& d) {
    PhiArgs a;
    a.nB = d.nB; a.nA = d.nA; a.r = d.r; a.row_width = d.row_width; a.col_width = d.col_width;
    a.rows = d.rows; a.cols = d.cols; a.mu = d.mu; a.wV = d.wV; a.wA0 = d.wA0; a.wA1 = d.wA1;
    a.a0_off = d.a0_off; a.a0_idx = d.a0_idx; a.a1_off = d.a1_off; a.a1_idx = d.a1_idx;
    a.row_off = d.row_off; a.row_rank = d.row_rank; a.row_beta = d.row_beta;
    a.col_off = d.col_off; a.col_rank = d.col_rank; a.col_a = d.col_a;
    a.row_inc = d.row_inc; a.row_inc_rank = d.row_inc_rank;
    a.col_inc = d.col_inc; a.col_inc_rank = d.col_inc_rank;
    return a;
}

// forward: row (s', beta); sum over incidences (i, s = s' + i) and a in the
// character block w(beta) - w(i) of mu[i][beta][a] * in[colbase(s) + local_a]
template <int B>
__global__ __launch_bounds__(256) void k_cyclic_fwd(PhiArgs a, const uint32_t* __restrict__ in,
                                                    uint32_t* __restrict__ out,
                                                    const uint32_t* __restrict__ scale, ModP m) {
    constexpr int RPB = 256 / B;
    const uint64_t row = (uint64_t)blockIdx.x * RPB + threadIdx.x / B;
    const int c = threadIdx.x % B;
    if (threadIdx.x / B >= RPB || row >= a.rows) return;
    const uint32_t beta = a.row_beta[row];
    const uint32_t bw = a.wA1[beta];
    const uint64_t s = a.row_rank[row];
    uint64_t lo = 0, hi = 0;
    for (uint32_t pos = 0; pos < a.row_width; ++pos) {
        const uint32_t code = a.row_inc[s * a.row_width + pos];
        const uint32_t i = code & 0x7fffffffu;
        const bool neg = code >> 31;
        const uint32_t q = (bw + a.r - (a.wV[i] % a.r)) % a.r;
        const uint32_t start = a.a0_off[q], end = a.a0_off[q + 1];
        const uint64_t base = a.col_off[a.row_inc_rank[s * a.row_width + pos]];
        const uint32_t* murow = a.mu + ((uint64_t)i * a.nB + beta) * a.nA;
        for (uint32_t idx = start; idx < end; ++idx) {
            uint32_t t = murow[a.a0_idx[idx]];
            if (t == 0) continue;
            if (neg) t = m.p - t;
            uint64_t prod = (uint64_t)t * in[(base + (idx - start)) * B + c];
            lo += (uint32_t)prod;
            hi += prod >> 32;
        }
    }
    uint32_t v = mod_split(hi, lo, m);
    if (scale) v = mulmod(v, scale[row], m);
    out[row * B + c] = v;
}

// transpose: column (s, a); sum over incidences (i, s - i) and beta in the
// character block w(a) + w(i) of mu[i][beta][a] * in[rowbase(s - i) + local_beta]
template <int B>
__global__ __launch_bounds__(256) void k_cyclic_trn(PhiArgs a, const uint32_t* __restrict__ in,
                                                    uint32_t* __restrict__ out,
                                                    const uint32_t* __restrict__ scale, ModP m) {
    constexpr int RPB = 256 / B;
    const uint64_t col = (uint64_t)blockIdx.x * RPB + threadIdx.x / B;
    const int c = threadIdx.x % B;
    if (threadIdx.x / B >= RPB || col >= a.cols) return;
    const uint32_t av = a.col_a[col];
    const uint32_t aw = a.wA0[av];
    const uint64_t s = a.col_rank[col];
    uint64_t lo = 0, hi = 0;
    for (uint32_t pos = 0; pos < a.col_width; ++pos) {
        const uint32_t code = a.col_inc[s * a.col_width + pos];
        const uint32_t i = code & 0x7fffffffu;
        const bool neg = code >> 31;
        const uint32_t q = (aw + a.wV[i]) % a.r;
        const uint32_t start = a.a1_off[q], end = a.a1_off[q + 1];
        const uint64_t base = a.row_off[a.col_inc_rank[s * a.col_width + pos]];
        for (uint32_t idx = start; idx < end; ++idx) {
            const uint32_t beta = a.a1_idx[idx];
            uint32_t t = a.mu[((uint64_t)i * a.nB + beta) * a.nA + av];
            if (t == 0) continue;
            if (neg) t = m.p - t;
            uint64_t prod = (uint64_t)t * in[(base + (idx - start)) * B + c];
            lo += (uint32_t)prod;
            hi += prod >> 32;
        }
    }
    uint32_t v = mod_split(hi, lo, m);
    if (scale) v = mulmod(v, scale[col], m);
    out[col * B + c] = v;
}

void launch_cyclic_fwd(const PhiDevice& d, const uint32_t* in, uint32_t* out,
                       const uint32_t* scale, uint32_t b, const ModP& m, hipStream_t st) {
    if (d.rows == 0) return;
    PhiArgs a = phi_args(d);
#define PGW_CALL(BB)                                                                         \
    hipLaunchKernelGGL(k_cyclic_fwd<BB>, dim3(grid_for(d.rows, 256 / BB)), dim3(256), 0, st, \
                       a, in, out, scale, m)
    PGW_SWITCH_B(b, PGW_CALL)
#undef PGW_CALL
    PGW_CHECK_LAUNCH("cyclic forward sweep");
}

void launch_cyclic_trn(const PhiDevice& d, const uint32_t* in, uint32_t* out,
                       const uint32_t* scale, uint32_t b, const ModP& m, hipStream_t st) {
    if (d.cols == 0) return;
    PhiArgs a = phi_args(d);
#define PGW_CALL(BB)                                                                         \
    hipLaunchKernelGGL(k_cyclic_trn<BB>, dim3(grid_for(d.cols, 256 / BB)), dim3(256), 0, st, \
                       a, in, out, scale, m)
    PGW_SWITCH_B(b, PGW_CALL)
#undef PGW_CALL
    PGW_CHECK_LAUNCH("cyclic transpose sweep");
}

// ===========================================================================
// elimination bridge (cuda/cyclic_cuda_helpers.h:718-895)
// ===========================================================================

// y[s,a] = sum_kappa Bk[a][kappa] z[s,kappa] - sum_beta R[a][beta] t[s,beta]
template <int B>
__global__ __launch_bounds__(256) void k_elim_y(
    uint64_t y_rows, const uint32_t* __restrict__ y_rank, const uint16_t* __restrict__ y_a,
    const uint32_t* __restrict__ z_off, const uint16_t* __restrict__ z_kappa,
    const uint32_t* __restrict__ t_off, const uint16_t* __restrict__ t_beta,
    const uint32_t* __restrict__ right_inv, const uint32_t* __restrict__ kbasis, uint32_t n_v,
    uint32_t kdim, const uint32_t* __restrict__ z, const uint32_t* __restrict__ t,
    uint32_t* __restrict__ y, ModP m) {
    constexpr int RPB = 256 / B;
    const uint64_t row = (uint64_t)blockIdx.x * RPB + threadIdx.x / B;
    const int c = threadIdx.x % B;
    if (threadIdx.x / B >= RPB || row >= y_rows) return;
    const uint32_t s = y_rank[row], a = y_a[row];
    uint64_t plo = 0, phi = 0, mlo = 0, mhi = 0;
    for (uint32_t idx = z_off[s]; idx < z_off[s + 1]; ++idx) {
        uint64_t prod = (uint64_t)kbasis[(uint64_t)a * kdim + z_kappa[idx]] * z[(uint64_t)idx * B + c];
        plo += (uint32_t)prod;
        phi += prod >> 32;
    }
    for (uint32_t idx = t_off[s]; idx < t_off[s + 1]; ++idx) {
        uint64_t prod = (uint64_t)right_inv[(uint64_t)a * n_v + t_beta[idx]] * t[(uint64_t)idx * B + c];
        mlo += (uint32_t)prod;
        mhi += prod >> 32;
    }
    uint32_t pv = mod_split(phi, plo, m), mv = mod_split(mhi, mlo, m);
    y[row * B + c] = addmod(pv, mv ? m.p - mv : 0u, m.p);
}

// t_adj[s,beta] = sum_a R[a][beta] h[s,a]
template <int B>
__global__ __launch_bounds__(256) void k_elim_rt_h(
    uint64_t t_rows, const uint32_t* __restrict__ t_rank, const uint16_t* __restrict__ t_beta,
    const uint32_t* __restrict__ y_off, const uint16_t* __restrict__ y_a,
    const uint32_t* __restrict__ right_inv, uint32_t n_v, const uint32_t* __restrict__ h,
    uint32_t* __restrict__ t_adj, ModP m) {
    constexpr int RPB = 256 / B;
    const uint64_t row = (uint64_t)blockIdx.x * RPB + threadIdx.x / B;
    const int c = threadIdx.x % B;
    if (threadIdx.x / B >= RPB || row >= t_rows) return;
    const uint32_t s = t_rank[row], beta = t_beta[row];
    uint64_t lo = 0, hi = 0;
    for (uint32_t idx = y_off[s]; idx < y_off[s + 1]; ++idx) {
        uint64_t prod = (uint64_t)right_inv[(uint64_t)y_a[idx] * n_v + beta] * h[(uint64_t)idx * B + c];
        lo += (uint32_t)prod;
        hi += prod >> 32;
    }
    t_adj[row * B + c] = mod_split(hi, lo, m);
}

// z_adj[s,kappa] = scale * sum_a Bk[a][kappa] h[s,a]
template <int B>
__global__ __launch_bounds__(256) void k_elim_bt_h(
    uint64_t z_rows, const uint32_t* __restrict__ z_rank, const uint16_t* __restrict__ z_kappa,
    const uint32_t* __restrict__ y_off, const uint16_t* __restrict__ y_a,
    const uint32_t* __restrict__ kbasis, uint32_t kdim, const uint32_t* __restrict__ h,
    uint32_t* __restrict__ z_adj, const uint32_t* __restrict__ scale, ModP m) {
    constexpr int RPB = 256 / B;
    const uint64_t row = (uint64_t)blockIdx.x * RPB + threadIdx.x / B;
    const int c = threadIdx.x % B;
    if (threadIdx.x / B >= RPB || row >= z_rows) return;
    const uint32_t s = z_rank[row], kappa = z_kappa[row];
    uint64_t lo = 0, hi = 0;
    for (uint32_t idx = y_off[s]; idx < y_off[s + 1]; ++idx) {
        uint64_t prod = (uint64_t)kbasis[(uint64_t)y_a[idx] * kdim + kappa] * h[(uint64_t)idx * B + c];
        lo += (uint32_t)prod;
        hi += prod >> 32;
    }
    uint32_t v = mod_split(hi, lo, m);
    if (scale) v = mulmod(v, scale[row], m);
    z_adj[row * B + c] = v;
}

void launch_elim_y(const PhiDevice& dm, const PhiDevice& d2, const uint32_t* z_off,
                   const uint16_t* z_kappa, const uint32_t* right_inv, const uint32_t* kbasis,
                   uint32_t n_v, uint32_t kdim, const uint32_t* z, const uint32_t* t, uint32_t* y,
                   uint32_t b, const ModP& m, hipStream_t st) {
    if (d2.cols == 0) return;
#define PGW_CALL(BB)                                                                           \
    hipLaunchKernelGGL(k_elim_y<BB>, dim3(grid_for(d2.cols, 256 / BB)), dim3(256), 0, st,      \
                       d2.cols, d2.col_rank, d2.col_a, z_off, z_kappa, dm.row_off, dm.row_beta, \
                       right_inv, kbasis, n_v, kdim, z, t, y, m)
    PGW_SWITCH_B(b, PGW_CALL)
#undef PGW_CALL
    PGW_CHECK_LAUNCH("elimination y");
}

void launch_elim_rt_h(const PhiDevice& dm, const PhiDevice& d2, const uint32_t* right_inv,
                      uint32_t n_v, const uint32_t* h, uint32_t* t_adj, uint32_t b,
                      const ModP& m, hipStream_t st) {
    if (dm.rows == 0) return;
#define PGW_CALL(BB)                                                                          \
    hipLaunchKernelGGL(k_elim_rt_h<BB>, dim3(grid_for(dm.rows, 256 / BB)), dim3(256), 0, st,  \
                       dm.rows, dm.row_rank, dm.row_beta, d2.col_off, d2.col_a, right_inv, n_v, \
                       h, t_adj, m)
    PGW_SWITCH_B(b, PGW_CALL)
#undef PGW_CALL
    PGW_CHECK_LAUNCH("elimination R^T h");
}

void launch_elim_bt_h(const PhiDevice& d2, uint64_t z_count, const uint32_t* z_rank,
                      const uint16_t* z_kappa, const uint32_t* kbasis, uint32_t kdim,
                      const uint32_t* h, uint32_t* z_adj, const uint32_t* scale, uint32_t b,
                      const ModP& m, hipStream_t st) {
    if (z_count == 0) return;
#define PGW_CALL(BB)                                                                          \
    hipLaunchKernelGGL(k_elim_bt_h<BB>, dim3(grid_for(z_count, 256 / BB)), dim3(256), 0, st,  \
                       z_count, z_rank, z_kappa, d2.col_off, d2.col_a, kbasis, kdim, h, z_adj, \
                       scale, m)
    PGW_SWITCH_B(b, PGW_CALL)
#undef PGW_CALL
    PGW_CHECK_LAUNCH("elimination B^T h");
}

// ===========================================================================
// elementwise: row scaling and the target preconditioner P^T Dr P
// ===========================================================================

__global__ __launch_bounds__(256) void k_scale_rows(const uint32_t* __restrict__ in,
                                                    const uint32_t* __restrict__ s,
                                                    uint32_t* __restrict__ out, uint64_t total,
                                                    uint32_t b, ModP m) {
    for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * 256)
        out[e] = mulmod(in[e], s[e / b], m);
}

void launch_scale_rows(const uint32_t* in, const uint32_t* s, uint32_t* out, uint64_t rows,
                       uint32_t b, const ModP& m, hipStream_t st) {
    const uint64_t total = rows * b;
    if (total == 0) return;
    unsigned g = (unsigned)std::min<uint64_t>(grid_for(total), 65536);
    hipLaunchKernelGGL(k_scale_rows, dim3(g), dim3(256), 0, st, in, s, out, total, b, m);
    PGW_CHECK_LAUNCH("scale rows");
}

// ===========================================================================
// deformation variants: dense columns beside the eliminated cyclic operator
// (cyclic_dense_columns_*_kernel, cyclic_insert_complement_scaled_kernel,
// cyclic_gather_complement_kernel: cuda/cyclic_cuda_helpers.h:1210-1351)
// ===========================================================================

// base position of kept column j (the base columns in order, drop0 < drop1 left out)
__device__ __forceinline__ uint64_t kept_to_base(uint64_t j, uint32_t drop0, uint32_t drop1) {
    if (j >= drop0) ++j;
    if (j >= drop1) ++j;
    return j;
}

__global__ __launch_bounds__(256) void k_complement_scale(const uint32_t* __restrict__ in,
                                                          const uint32_t* __restrict__ dc,
                                                          uint32_t* __restrict__ scaled, uint64_t base,
                                                          uint32_t drop0, uint32_t drop1, uint64_t total,
                                                          uint32_t b, ModP m) {
    for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * 256) {
        const uint64_t r = e / b, c = e % b;
        uint64_t src;  // input row feeding scaled row r
        if (r < base) {
            if (r == drop0 || r == drop1) {
                scaled[e] = 0;
                continue;
            }
            src = r - (r > drop0 ? 1 : 0) - (r > drop1 ? 1 : 0);
        } else {
            src = r - 2;  // dense-column inputs follow the base - 2 kept columns
        }
        scaled[e] = mulmod(in[src * b + c], dc[src], m);
    }
}

void launch_complement_scale(const uint32_t* in, const uint32_t* dc, uint32_t* scaled,
                             uint64_t base, uint32_t drop0, uint32_t drop1, uint32_t dcols,
                             uint32_t b, const ModP& m, hipStream_t st) {
    const uint64_t total = (base + dcols) * b;
    if (total == 0) return;
    unsigned g = (unsigned)std::min<uint64_t>(grid_for(total), 65536);
    hipLaunchKernelGGL(k_complement_scale, dim3(g), dim3(256), 0, st, in, dc, scaled, base, drop0,
                       drop1, total, b, m);
    PGW_CHECK_LAUNCH("complement scale");
}

__global__ __launch_bounds__(256) void k_complement_gather(const uint32_t* __restrict__ full,
                                                           uint32_t* __restrict__ out, uint64_t total,
                                                           uint32_t drop0, uint32_t drop1, uint32_t b) {
    for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * 256) {
        const uint64_t j = e / b, c = e % b;
        out[e] = full[kept_to_base(j, drop0, drop1) * b + c];
    }
}

void launch_complement_gather(const uint32_t* full, uint32_t* out, uint64_t base, uint32_t drop0,
                              uint32_t drop1, uint32_t b, hipStream_t st) {
    const uint64_t total = (base - 2) * b;
    if (total == 0) return;
    unsigned g = (unsigned)std::min<uint64_t>(grid_for(total), 65536);
    hipLaunchKernelGGL(k_complement_gather, dim3(g), dim3(256), 0, st, full, out, total, drop0, drop1, b);
    PGW_CHECK_LAUNCH("complement gather");
}

// one thread per target entry: sum_j dense[r][j] * extra[j][c], each product
// reduced (any dcols), then added to or subtracted from the target entry
__global__ __launch_bounds__(256) void k_dense_cols_axpy(uint32_t* __restrict__ target,
                                                         const uint32_t* __restrict__ dense,
                                                         const uint32_t* __restrict__ extra,
                                                         uint64_t total, uint32_t dcols, uint32_t b,
                                                         int subtract, ModP m) {
    for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * 256) {
        const uint64_t r = e / b, c = e % b;
        uint64_t sum = 0;  // dcols terms below 2^31 each
        for (uint32_t j = 0; j < dcols; ++j)
            sum += mulmod(dense[r * dcols + j], extra[(uint64_t)j * b + c], m);
        const uint32_t t = mod_u64(sum, m);
        const uint32_t x = target[e];
        target[e] = subtract ? addmod(x, t ? m.p - t : 0u, m.p) : addmod(x, t, m.p);
    }
}

void launch_dense_cols_axpy(uint32_t* target, const uint32_t* dense, const uint32_t* extra,
                            uint64_t rows, uint32_t dcols, uint32_t b, bool subtract,
                            const ModP& m, hipStream_t st) {
    const uint64_t total = rows * b;
    if (total == 0) return;
    unsigned g = (unsigned)std::min<uint64_t>(grid_for(total), 65536);
    hipLaunchKernelGGL(k_dense_cols_axpy, dim3(g), dim3(256), 0, st, target, dense, extra, total,
                       dcols, b, subtract ? 1 : 0, m);
    PGW_CHECK_LAUNCH("dense columns axpy");
}

// dense^T y in two passes and without atomics: each of at most kDenseTBlocks
// workgroups sums its rows (a grid-stride walk) for every (j, c) pair with
// reduced products in u64 (exact for any row count below 2^32), the second
// pass adds the partials, reduces once, signs and scales.
constexpr uint32_t kDenseTBlocks = 256;

size_t dense_cols_scratch_bytes(uint32_t dcols, uint32_t b) {
    return (size_t)kDenseTBlocks * dcols * b * sizeof(uint64_t);
}

__global__ __launch_bounds__(256) void k_dense_cols_t_partial(const uint32_t* __restrict__ dense,
                                                              const uint32_t* __restrict__ y,
                                                              uint64_t rows, uint32_t dcols, uint32_t b,
                                                              uint64_t* __restrict__ partial, ModP m) {
    __shared__ uint64_t s[256];
    const uint32_t np = dcols * b;  // (j, c) pairs, np <= 256
    const uint32_t rl = 256 / np;   // rows walked side by side
    const uint32_t t = threadIdx.x, pair = t % np, lane = t / np;
    uint64_t acc = 0;
    if (lane < rl) {
        const uint32_t j = pair / b, c = pair % b;
        for (uint64_t r = (uint64_t)blockIdx.x * rl + lane; r < rows; r += (uint64_t)gridDim.x * rl)
            acc += mulmod(dense[r * dcols + j], y[r * b + c], m);
    }
    s[t] = acc;
    __syncthreads();
    if (t < np) {
        uint64_t sum = 0;
        for (uint32_t q = 0; q < rl; ++q) sum += s[q * np + t];
        partial[(uint64_t)blockIdx.x * np + t] = sum;
    }
}

__global__ void k_dense_cols_t_finish(const uint64_t* __restrict__ partial, uint32_t blocks,
                                      uint32_t dcols, uint32_t b, const uint32_t* __restrict__ dc,
                                      uint64_t lead, uint32_t* __restrict__ out, int subtract, ModP m) {
    const uint32_t np = dcols * b, t = threadIdx.x;
    if (t >= np) return;
    uint64_t sum = 0;  // blocks * (rows / blocks) * 2^31 < 2^63
    for (uint32_t q = 0; q < blocks; ++q) sum += partial[(uint64_t)q * np + t];
    uint32_t v = mod_u64(sum, m);
    if (subtract && v) v = m.p - v;
    const uint32_t j = t / b, c = t % b;
    out[(lead + j) * b + c] = mulmod(v, dc[lead + j], m);
}

void launch_dense_cols_transpose(const uint32_t* dense, const uint32_t* y, uint64_t rows,
                                 uint32_t dcols, uint32_t b, const uint32_t* dc, uint64_t lead,
                                 uint32_t* out, bool subtract, uint64_t* scratch, const ModP& m,
                                 hipStream_t st) {
    const uint32_t np = dcols * b;
    require(np >= 1 && np <= 256, "dense columns x block width must be at most 256");
    const uint32_t rl = 256 / np;
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(kDenseTBlocks, (rows + rl - 1) / rl));
    hipLaunchKernelGGL(k_dense_cols_t_partial, dim3(blocks), dim3(256), 0, st, dense, y, rows, dcols,
                       b, scratch, m);
    hipLaunchKernelGGL(k_dense_cols_t_finish, dim3(1), dim3(256), 0, st, scratch, blocks, dcols, b,
                       dc, lead, out, subtract ? 1 : 0, m);
    PGW_CHECK_LAUNCH("dense columns transpose");
}

// v mod p for v < 2^63.  M31 (p = 2^31 - 1): 2^31 == 1, so two folds of the
// high bits onto the low 31 bits leave v' < 2^31 + 2 and one conditional
// subtraction finishes; otherwise 64-bit Barrett.
template <bool M31>
__device__ __forceinline__ uint32_t red63(uint64_t v, const ModP& m) {
    if constexpr (M31) {
        uint64_t r = (v & 0x7fffffffULL) + (v >> 31);  // < 2^32 + 2^31
        r = (r & 0x7fffffffULL) + (r >> 31);           // < 2^31 + 3
        return (uint32_t)(r >= 0x7fffffffULL ? r - 0x7fffffffULL : r);
    } else {
        return mod_u64(v, m);
    }
}

// one round of paired 2x2 maps; forward [[1,a],[b,1+ab]], transpose [[1,b],[a,1+ab]]
// (cuda/cyclic_cuda_helpers.h:109-143).  `mid` fuses forward, Dr and transpose
// of the same round: the 2x2 block P_r^T diag(dr_f, dr_s) P_r on each pair.
// Every sum below is u*x + v*y with residues below 2^31, so below 2^63.
template <int B, int MODE, bool M31>  // MODE 0 forward, 1 transpose, 2 fused middle
__global__ __launch_bounds__(256) void k_rowmix(const pgw_rowmix_pair* __restrict__ pairs,
                                                uint64_t count, uint32_t* __restrict__ block,
                                                const uint32_t* __restrict__ dr, ModP m) {
    constexpr int RPB = 256 / B;
    const uint64_t q = (uint64_t)blockIdx.x * RPB + threadIdx.x / B;
    const int c = threadIdx.x % B;
    if (threadIdx.x / B >= RPB || q >= count) return;
    const pgw_rowmix_pair pr = pairs[q];
    const uint32_t p = m.p;
    const uint32_t opab = addmod(1u, red63<M31>((uint64_t)pr.a * pr.b, m), p);
    uint32_t* fx = block + (uint64_t)pr.first * B + c;
    uint32_t* sy = block + (uint64_t)pr.second * B + c;
    uint64_t x = *fx, y = *sy;
    uint32_t nx, ny;
    if (MODE == 1) {
        nx = red63<M31>(x + (uint64_t)pr.b * y, m);
        ny = red63<M31>((uint64_t)pr.a * x + (uint64_t)opab * y, m);
    } else {
        nx = red63<M31>(x + (uint64_t)pr.a * y, m);
        ny = red63<M31>((uint64_t)pr.b * x + (uint64_t)opab * y, m);
        if (MODE == 2) {
            const uint64_t sx = red63<M31>((uint64_t)nx * dr[pr.first], m);
            const uint64_t sy2 = red63<M31>((uint64_t)ny * dr[pr.second], m);
            nx = red63<M31>(sx + (uint64_t)pr.b * sy2, m);
            ny = red63<M31>((uint64_t)pr.a * sx + (uint64_t)opab * sy2, m);
        }
    }
    *fx = nx;
    *sy = ny;
}

// the row left out of the last round's pairs (odd row count) is fixed by that
// round's map, so P_last^T Dr P_last only scales it by its Dr entry
__global__ void k_scale_one_row(uint32_t* __restrict__ block, uint64_t row,
                                const uint32_t* __restrict__ dr, uint32_t b, ModP m) {
    if (threadIdx.x < b) block[row * b + threadIdx.x] = mulmod(block[row * b + threadIdx.x], dr[row], m);
}

template <int B, bool M31>
static void target_precond_bm(uint32_t* block, const pgw_rowmix_pair* pairs, uint64_t pc,
                              uint32_t rounds, const uint32_t* dr, int64_t unpaired_row,
                              const ModP& m, hipStream_t st) {
    dim3 grid(grid_for(pc, 256 / B)), blk(256);
    for (uint32_t r = 0; r + 1 < rounds; ++r)
        hipLaunchKernelGGL((k_rowmix<B, 0, M31>), grid, blk, 0, st, pairs + (uint64_t)r * pc, pc,
                           block, dr, m);
    hipLaunchKernelGGL((k_rowmix<B, 2, M31>), grid, blk, 0, st, pairs + (uint64_t)(rounds - 1) * pc,
                       pc, block, dr, m);
    if (unpaired_row >= 0)
        hipLaunchKernelGGL(k_scale_one_row, dim3(1), dim3(64), 0, st, block, (uint64_t)unpaired_row,
                           dr, (uint32_t)B, m);
    for (uint32_t r = rounds - 1; r-- > 0;)
        hipLaunchKernelGGL((k_rowmix<B, 1, M31>), grid, blk, 0, st, pairs + (uint64_t)r * pc, pc,
                           block, dr, m);
    PGW_CHECK_LAUNCH("row mixing");
}

template <int B>
static void target_precond_b(uint32_t* block, const pgw_rowmix_pair* pairs, uint64_t pc,
                             uint32_t rounds, const uint32_t* dr, int64_t unpaired_row,
                             const ModP& m, hipStream_t st) {
    if (m.p == 0x7fffffffu) target_precond_bm<B, true>(block, pairs, pc, rounds, dr, unpaired_row, m, st);
    else target_precond_bm<B, false>(block, pairs, pc, rounds, dr, unpaired_row, m, st);
}

void launch_target_precond(uint32_t* block, const pgw_rowmix_pair* pairs, uint64_t pair_count,
                           uint32_t rounds, const uint32_t* dr, int64_t unpaired_row,
                           uint64_t rows, uint32_t b, const ModP& m, hipStream_t st) {
    if (rounds == 0 || pair_count == 0) {
        launch_scale_rows(block, dr, block, rows, b, m, st);
        return;
    }
    require(rows == 2 * pair_count || rows == 2 * pair_count + 1, "row-mix plan does not cover the target");
    require((rows % 2 == 1) == (unpaired_row >= 0), "unpaired row does not match the row count");
#define PGW_CALL(BB) target_precond_b<BB>(block, pairs, pair_count, rounds, dr, unpaired_row, m, st)
    PGW_SWITCH_B(b, PGW_CALL)
#undef PGW_CALL
}

// ===========================================================================
// Gram projection: packed upper triangle of X^T Y (and Y^T Y or X^T X)
// ===========================================================================
//
// Each workgroup stages 64-row chunks of X and Y in LDS; thread (i,j) keeps
// 64-bit low/high word sums of the products (each < 2^62), so partial sums are
// exact for any N < 2^32 rows; a second kernel adds the partials and reduces.

constexpr int kGramRows = 64;
constexpr unsigned kGramGrid = 1024;

size_t gram_scratch_bytes(uint32_t b) {
    return (size_t)kGramGrid * 2 * 256 * 2 * sizeof(uint64_t);
}

template <int B>
__global__ __launch_bounds__(256) void k_gram_partial(const uint32_t* __restrict__ x,
                                                      const uint32_t* __restrict__ y, uint64_t rows,
                                                      int mode, uint64_t* __restrict__ part) {
    __shared__ uint32_t sx[kGramRows * B];
    __shared__ uint32_t sy[kGramRows * B];
    const int t = threadIdx.x;
    const int i = t / 16, j = t % 16;
    const bool active = i < B && j < B && i <= j;
    uint64_t lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0;
    for (uint64_t r0 = (uint64_t)blockIdx.x * kGramRows; r0 < rows;
         r0 += (uint64_t)gridDim.x * kGramRows) {
        const uint64_t nrow = min((uint64_t)kGramRows, rows - r0);
        __syncthreads();
        for (int e = t; e < kGramRows * B; e += 256) {
            bool ok = (uint64_t)(e / B) < nrow;
            sx[e] = ok ? x[r0 * B + e] : 0u;
            sy[e] = ok ? y[r0 * B + e] : 0u;
        }
        __syncthreads();
        if (active) {
            for (int r = 0; r < kGramRows; ++r) {
                uint64_t p0 = (uint64_t)sx[r * B + i] * sy[r * B + j];
                lo0 += (uint32_t)p0;
                hi0 += p0 >> 32;
                if (mode) {
                    uint64_t p1 = mode == 1 ? (uint64_t)sy[r * B + i] * sy[r * B + j]
                                            : (uint64_t)sx[r * B + i] * sx[r * B + j];
                    lo1 += (uint32_t)p1;
                    hi1 += p1 >> 32;
                }
            }
        }
    }
    if (active) {
        uint64_t* o = part + ((uint64_t)blockIdx.x * 2 * 256 + (uint64_t)(i * 16 + j)) * 2;
        o[0] = lo0;
        o[1] = hi0;
        o[512] = lo1;
        o[513] = hi1;
    }
}

// With `pos` set, out0/out1 are offset by ut * (*pos + pos_off) entries: the
// sequence position is read on the device, so a captured graph of a whole
// round can be replayed unchanged.
__global__ __launch_bounds__(256) void k_gram_finish(const uint64_t* __restrict__ part,
                                                     unsigned nblocks, uint32_t b, int mode,
                                                     uint32_t* __restrict__ out0,
                                                     uint32_t* __restrict__ out1, ModP m,
                                                     const uint32_t* __restrict__ pos,
                                                     uint32_t pos_off) {
    const int t = threadIdx.x;
    const int i = t / 16, j = t % 16;
    if (i >= (int)b || j >= (int)b || i > j) return;
    if (pos) {
        const uint64_t shift = (uint64_t)(b * (b + 1) / 2) * (*pos + pos_off);
        out0 += shift;
        if (out1) out1 += shift;
    }
    const int ut = i * (2 * (int)b - i + 1) / 2 + (j - i);
    for (int mat = 0; mat < (mode ? 2 : 1); ++mat) {
        uint64_t lo = 0, hi = 0;
        for (unsigned g = 0; g < nblocks; ++g) {
            const uint64_t* o = part + (((uint64_t)g * 2 + mat) * 256 + t) * 2;
            lo += o[0];
            hi += o[1];
        }
        (mat ? out1 : out0)[ut] = mod_split(hi, lo, m);
    }
}

__global__ void k_advance_pos(uint32_t* pos, uint32_t by) {
    if (threadIdx.x == 0) *pos += by;
}

void launch_advance_pos(uint32_t* pos, uint32_t by, hipStream_t st) {
    hipLaunchKernelGGL(k_advance_pos, dim3(1), dim3(64), 0, st, pos, by);
    PGW_CHECK_LAUNCH("advance position");
}

void launch_gram(const uint32_t* x, const uint32_t* y, uint64_t rows, uint32_t b, int mode,
                 uint32_t* out0, uint32_t* out1, uint64_t* scratch, const ModP& m,
                 hipStream_t st, const uint32_t* pos, uint32_t pos_off) {
    // low words: rows * 2^32 < 2^64 requires rows < 2^32
    require(rows < (1ULL << 32), "Gram projection supports fewer than 2^32 rows");
    unsigned grid = (unsigned)std::min<uint64_t>(kGramGrid, std::max<uint64_t>(1, (rows + kGramRows - 1) / kGramRows));
#define PGW_CALL(BB)                                                                           \
    hipLaunchKernelGGL(k_gram_partial<BB>, dim3(grid), dim3(256), 0, st, x, y, rows, mode, scratch)
    PGW_SWITCH_B(b, PGW_CALL)
#undef PGW_CALL
    hipLaunchKernelGGL(k_gram_finish, dim3(1), dim3(256), 0, st, scratch, grid, b, mode, out0,
                       out1, m, pos, pos_off);
    PGW_CHECK_LAUNCH("gram");
}

// ===========================================================================
// kernel-vector generator slice (cuprym_cyclic_kernel_vectors.cu:1766-1790)
// ===========================================================================

template <int B>
__global__ __launch_bounds__(256) void k_accumulate_generator(const uint32_t* __restrict__ Bm,
                                                              const uint32_t* __restrict__ G,
                                                              uint32_t* __restrict__ D,
                                                              uint64_t rows, ModP m) {
    constexpr int RPB = 256 / B;
    const uint64_t row = (uint64_t)blockIdx.x * RPB + threadIdx.x / B;
    const int c = threadIdx.x % B;
    if (threadIdx.x / B >= RPB || row >= rows) return;
    uint64_t lo = D[row * B + c], hi = 0;
#pragma unroll
    for (int i = 0; i < B; ++i) {
        uint64_t prod = (uint64_t)Bm[row * B + i] * G[c * B + i];
        lo += (uint32_t)prod;
        hi += prod >> 32;
    }
    D[row * B + c] = mod_split(hi, lo, m);
}

void launch_accumulate_generator(const uint32_t* Bm, const uint32_t* G, uint32_t* D, uint64_t rows,
                                 uint32_t b, const ModP& m, hipStream_t st) {
    if (rows == 0) return;
#define PGW_CALL(BB)                                                                          \
    hipLaunchKernelGGL(k_accumulate_generator<BB>, dim3(grid_for(rows, 256 / BB)), dim3(256), \
                       0, st, Bm, G, D, rows, m)
    PGW_SWITCH_B(b, PGW_CALL)
#undef PGW_CALL
    PGW_CHECK_LAUNCH("accumulate generator");
}

}  // namespace pgw

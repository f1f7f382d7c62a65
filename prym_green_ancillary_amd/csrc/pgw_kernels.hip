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

#include <cstdlib>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <utility>

#include "pgw_fold.h"
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
// explicit sparse maps (the cyclic operator's chain, pgw_sparse.cpp)
// ===========================================================================
//
// out (rows x B) = M in (cols x B) mod p for a sliced-ELL map: the entries of
// slice s (kSellHeight consecutive rows) are interleaved row by row, so that
// the lanes of a wavefront that work on consecutive rows read consecutive
// entries.  A row is handled by B / CW column groups (a 16-byte load gathers
// four columns) times LANES lanes that split its entries (lane l takes entries
// l, l + LANES, ...) and add their sums with cross-lane shuffles.
// Every entry is a dependent pair of loads (the entry, then the input row it
// names), so a lane takes its entries U at a time: U entry loads, then U
// gathers, all in flight together before the first product.  The loads carry
// no branch (a branch around a load makes the compiler wait for it at the
// join): an entry past the row's slice width re-reads the last entry of the
// row and is then zeroed, so its gather reads input row 0, which exists
// whenever the slice holds any entry.
// SMALLP (p < 2^16): products < 2^32, summed in u64; otherwise products
// < 2^62 summed as 64-bit low/high word sums.  Exact for rows of any length
// below 2^32.
// Workgroup b of a launch of nb runs on XCD b % 8 (the hardware deals
// workgroups to the eight XCDs in turn).  With `contiguous`, blocks are
// renumbered so that each XCD takes a contiguous range of the rows, whose
// gathers share input lines in that XCD's L2, instead of every eighth block:
// a bijection on [0, nb).  It pays while a map's input is within reach of the
// L2s (g24's maps, 323 k rows: 55.7 against 59.1 us per apply); on larger maps
// the dealt order, all XCDs reading one region at a time, shares lines through
// the memory-side cache instead (g26: 211.9 against 214.5 us, g28: 1147
// against 1182 us), so launches of more than kSellXcdRows rows keep it.
constexpr uint64_t kSellXcdRows = 1ULL << 19;
__device__ __forceinline__ uint64_t xcd_block(uint32_t b, uint32_t nb, bool contiguous) {
    if (!contiguous) return b;
    const uint32_t x = b % 8, q = nb / 8, r = nb % 8;
    return (uint64_t)x * q + (x < r ? x : r) + b / 8;
}

// The sums of one (row, column group) of a sparse map: lo (and hi) as
// described above, complete in the row's lane 0 after the shuffles.
// FORM (SellDevice::form): 0 (column, coefficient) pairs; 1 one u32 per
// entry, column | coefficient << pack_bits; 2 per slice, either such u32
// words holding the column's offset from the slice's base column
// (sbase[s] = base | 1 << 31) or pairs (sbase[s] = 0), slice[s] counting u32
// words.  A slice is kSellHeight rows, so a wavefront never straddles two and
// its branch on the slice's form is uniform.
template <int B, int CW, int LANES, int U, bool SMALLP, int FORM>
__device__ __forceinline__ void sell_row_sums(const uint64_t* __restrict__ slice, const uint32_t* __restrict__ ent,
                                              const uint32_t* __restrict__ sbase,
                                              const uint32_t* __restrict__ in, uint32_t pack_bits, uint64_t row,
                                              uint32_t g, uint32_t lane, bool valid, uint64_t (&lo)[CW],
                                              uint64_t (&hi)[CW]) {
#pragma unroll
    for (int q = 0; q < CW; ++q) lo[q] = hi[q] = 0;
    if (valid) {
        const uint64_t sl = row / kSellHeight;
        const uint32_t sb = FORM == 2 ? sbase[sl] : 0u;
        const bool words = FORM == 1 || (FORM == 2 && (sb >> 31) != 0);  // one u32 per entry
        const uint32_t base = sb & 0x7fffffffu;
        const uint64_t b0 = slice[sl];
        const uint64_t width = (slice[sl + 1] - b0) / kSellHeight / (FORM == 2 && !words ? 2 : 1);
        const uint64_t e0 = (FORM == 2 ? 0 : b0) + row % kSellHeight;
        const uint32_t* sent = FORM == 2 ? ent + b0 : ent;  // even word offsets: pairs stay 8-byte aligned
        const uint32_t col_mask = FORM != 0 ? (1u << pack_bits) - 1u : 0u;
        for (uint64_t k0 = lane; k0 < width; k0 += (uint64_t)LANES * U) {
            uint2 c[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t k = k0 + (uint64_t)u * LANES;
                const uint64_t e = e0 + (k < width ? k : width - 1) * kSellHeight;
                if (FORM != 0 && words) {
                    const uint32_t w = sent[e];
                    c[u] = make_uint2(base + (w & col_mask), w >> pack_bits);
                } else {
                    c[u] = reinterpret_cast<const uint2*>(sent)[e];
                }
                if (k >= width) c[u] = make_uint2(0u, 0u);
            }
            uint32_t xv[U][CW];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t* xr = in + (uint64_t)c[u].x * B + g * CW;
                if constexpr (CW == 4) {
                    const uint4 v = *reinterpret_cast<const uint4*>(xr);
                    xv[u][0] = v.x; xv[u][1] = v.y; xv[u][2] = v.z; xv[u][3] = v.w;
                } else {
#pragma unroll
                    for (int q = 0; q < CW; ++q) xv[u][q] = xr[q];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int q = 0; q < CW; ++q) {
                    const uint64_t prod = (uint64_t)c[u].y * xv[u][q];
                    if (SMALLP) {
                        lo[q] += prod;
                    } else {
                        lo[q] += (uint32_t)prod;
                        hi[q] += prod >> 32;
                    }
                }
        }
    }
    if constexpr (LANES > 1) {
        // the lanes of one (row, group) are adjacent: xor offsets below LANES stay inside it
#pragma unroll
        for (int off = LANES / 2; off > 0; off >>= 1)
#pragma unroll
            for (int q = 0; q < CW; ++q) {
                lo[q] += __shfl_xor(lo[q], off, 64);
                if (!SMALLP) hi[q] += __shfl_xor(hi[q], off, 64);
            }
    }
}

template <int B, int CW, int LANES, int U, bool SMALLP, int FORM>
__global__ __launch_bounds__(256) void k_sell_spmm(const uint64_t* __restrict__ slice,
                                                   const uint32_t* __restrict__ ent,
                                                   const uint32_t* __restrict__ sbase, uint64_t rows,
                                                   const uint32_t* __restrict__ in,
                                                   uint32_t* __restrict__ out, ModP m, uint32_t pack_bits,
                                                   const uint32_t* __restrict__ out_row,
                                                   uint32_t* __restrict__ bump) {
    constexpr int G = B / CW;
    constexpr int T = G * LANES;  // threads per row; LANES > 1 only when T divides 64
    const uint64_t tid = xcd_block(blockIdx.x, gridDim.x, rows <= kSellXcdRows) * 256 + threadIdx.x;
    const uint64_t row = tid / T;
    const uint32_t g = (uint32_t)((tid % T) / LANES), lane = (uint32_t)(tid % LANES);
    const bool valid = row < rows;
    // the sequence loop's position counter moves on here, in the round's
    // first launch, while no other kernel reads it
    if (bump && tid == 0) *bump += 4;
    uint64_t lo[CW], hi[CW];
    sell_row_sums<B, CW, LANES, U, SMALLP, FORM>(slice, ent, sbase, in, pack_bits, row, g, lane, valid, lo, hi);
    if (valid && lane == 0) {
        const uint64_t orow = out_row ? out_row[row] : row;
#pragma unroll
        for (int q = 0; q < CW; ++q)
            out[orow * B + g * CW + q] = SMALLP ? mod_u64(lo[q], m) : mod_split(hi[q], lo[q], m);
    }
}

// The last map of a cyclic apply at b = 4, p < 2^16, with the block
// projection of the sequence round folded in: besides out = M in it sums,
// over its rows r, the products of out[r] with `other`[r] (the apply's input
// block) and with itself, the two packed upper triangles the round pushes:
// mode 1 (other^T out, out^T out), mode 2 (out^T other, out^T out).  Every
// product is < 2^32 and is reduced mod p at once, so a wavefront adds 64 of
// them in 32 bits and a workgroup's four wavefronts stay below 2^24; the
// workgroup's sums land as k_gram_finish's partials
// part[(e * nblocks + block) * 2 + {0, 1}] = (sum, 0), e = mat * 10 + pair.
template <int LANES, int U, int FORM>
__global__ __launch_bounds__(256) void k_sell_spmm_gram(const uint64_t* __restrict__ slice,
                                                        const uint32_t* __restrict__ ent,
                                                        const uint32_t* __restrict__ sbase, uint64_t rows,
                                                        const uint32_t* __restrict__ in,
                                                        uint32_t* __restrict__ out, ModP m, uint32_t pack_bits,
                                                        const uint32_t* __restrict__ other, int mode,
                                                        uint64_t* __restrict__ part) {
    constexpr int B = 4, P = 10;
    const uint64_t tid = xcd_block(blockIdx.x, gridDim.x, rows <= kSellXcdRows) * 256 + threadIdx.x;
    const uint64_t row = tid / LANES;
    const uint32_t lane = (uint32_t)(tid % LANES);
    const bool valid = row < rows;
    const bool mine = valid && lane == 0;
    uint32_t x[B] = {0, 0, 0, 0};
    if (mine) {
        const uint4 v = *reinterpret_cast<const uint4*>(other + row * B);
        x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
    }
    uint64_t lo[B], hi[B];
    sell_row_sums<B, B, LANES, U, true, FORM>(slice, ent, sbase, in, pack_bits, row, 0, lane, valid, lo, hi);
    uint32_t o[B] = {0, 0, 0, 0};
    if (mine) {
#pragma unroll
        for (int q = 0; q < B; ++q) o[q] = mod_u64(lo[q], m);
        *reinterpret_cast<uint4*>(out + row * B) = make_uint4(o[0], o[1], o[2], o[3]);
    }
    // x mod p for x < 2^32 and p < 2^16: Barrett with floor(2^32 / p)
    const uint32_t mu = (uint32_t)(m.barrett >> 32);
    auto red = [&](uint32_t v) {
        const uint32_t q = __umulhi(v, mu);
        const uint32_t r = v - q * m.p;
        return r >= m.p ? r - m.p : r;
    };
    uint32_t sums[2 * P];
    int e = 0;
#pragma unroll
    for (int i = 0; i < B; ++i)
#pragma unroll
        for (int j = i; j < B; ++j, ++e) {
            sums[e] = red(mode == 1 ? x[i] * o[j] : o[i] * x[j]);
            sums[P + e] = red(o[i] * o[j]);
        }
    // wavefront sums on the device library's cross-lane (DPP) reduction; every
    // lane of the workgroup reaches this point
#pragma unroll
    for (int k = 0; k < 2 * P; ++k) sums[k] = __ockl_wfred_add_u32(sums[k]);
    __shared__ uint32_t s_sum[4][2 * P];
    const int wave = threadIdx.x / 64;
    if (threadIdx.x % 64 == 0)
#pragma unroll
        for (int k = 0; k < 2 * P; ++k) s_sum[wave][k] = sums[k];
    __syncthreads();
    if (threadIdx.x < 2 * P) {
        const uint32_t t = threadIdx.x;
        const uint64_t v = (uint64_t)s_sum[0][t] + s_sum[1][t] + s_sum[2][t] + s_sum[3][t];
        uint64_t* o2 = part + ((uint64_t)t * gridDim.x + blockIdx.x) * 2;
        o2[0] = v;
        o2[1] = 0;
    }
}

// entries a lane keeps in flight (k_sell_spmm's U)
constexpr int kSellBatch = 8;

template <int B, int CW, int LANES>
static void sell_launch(const SellDevice& M, const uint32_t* in, uint32_t* out, const ModP& m,
                        hipStream_t st, const uint32_t* out_row, uint32_t* bump) {
    const uint64_t threads = M.rows * (B / CW) * LANES;
    const dim3 grid(grid_for(threads)), blk(256);
#define PGW_SELL(SP, FORM)                                                                                \
    hipLaunchKernelGGL((k_sell_spmm<B, CW, LANES, kSellBatch, SP, FORM>), grid, blk, 0, st, M.slice, M.ent,   \
                       M.sbase, M.rows, in, out, m, M.pack_bits, out_row, bump)
    if (m.p < 65536u) {
        if (M.form == 2) PGW_SELL(true, 2); else if (M.form == 1) PGW_SELL(true, 1); else PGW_SELL(true, 0);
    } else {
        if (M.form == 2) PGW_SELL(false, 2); else if (M.form == 1) PGW_SELL(false, 1); else PGW_SELL(false, 0);
    }
#undef PGW_SELL
}

// lanes per row: with kSellBatch entries in flight per lane, one lane per row
// keeps the memory system busy from about 2^16 threads on (g22's 84 k rows:
// 37.4 us per apply with one lane, 37.7 with two, 51.4 with four; g24's 323 k
// rows: 69 / 75 / 83 us); fewer rows get more lanes, never more than the
// map's mean padded row width.  1 when B / CW is not a power of two (a row's
// threads must stay inside one wavefront)
template <int B, int CW>
static void sell_b(const SellDevice& M, const uint32_t* in, uint32_t* out, const ModP& m,
                   hipStream_t st, const uint32_t* out_row, uint32_t* bump) {
    constexpr int G = B / CW;
    int lanes = 1;
    if ((G & (G - 1)) == 0) {
        const double width = M.rows ? (double)M.padded / M.rows : 0.0;
        while (lanes < 8 && (double)M.rows * G * lanes < (double)(1 << 16) && 2.0 * lanes <= width)
            lanes *= 2;
    }
    if constexpr ((G & (G - 1)) == 0 && G <= 8) {
        if (lanes == 8) return sell_launch<B, CW, 8>(M, in, out, m, st, out_row, bump);
        if (lanes == 4) return sell_launch<B, CW, 4>(M, in, out, m, st, out_row, bump);
        if (lanes == 2) return sell_launch<B, CW, 2>(M, in, out, m, st, out_row, bump);
    }
    sell_launch<B, CW, 1>(M, in, out, m, st, out_row, bump);
}

void launch_sell_spmm(const SellDevice& M, const uint32_t* in, uint32_t* out, uint32_t b,
                      const ModP& m, hipStream_t st, const uint32_t* out_row, uint32_t* bump) {
    require(M.rows > 0 || bump == nullptr, "position update on an empty sparse map");
    if (M.rows == 0) return;
    require(M.rows < (1ULL << 40), "sparse map too large");
    // 16-byte gathers need 16-byte aligned rows
    const bool vec = b % 4 == 0 && (reinterpret_cast<uintptr_t>(in) % 16) == 0;
    switch (b) {
        case 1: sell_b<1, 1>(M, in, out, m, st, out_row, bump); break;
        case 2: sell_b<2, 2>(M, in, out, m, st, out_row, bump); break;
        case 3: sell_b<3, 3>(M, in, out, m, st, out_row, bump); break;
        case 4: if (vec) sell_b<4, 4>(M, in, out, m, st, out_row, bump); else sell_b<4, 1>(M, in, out, m, st, out_row, bump); break;
        case 5: sell_b<5, 1>(M, in, out, m, st, out_row, bump); break;
        case 6: sell_b<6, 2>(M, in, out, m, st, out_row, bump); break;
        case 7: sell_b<7, 1>(M, in, out, m, st, out_row, bump); break;
        case 8: if (vec) sell_b<8, 4>(M, in, out, m, st, out_row, bump); else sell_b<8, 1>(M, in, out, m, st, out_row, bump); break;
        case 9: sell_b<9, 3>(M, in, out, m, st, out_row, bump); break;
        case 10: sell_b<10, 2>(M, in, out, m, st, out_row, bump); break;
        case 11: sell_b<11, 1>(M, in, out, m, st, out_row, bump); break;
        case 12: if (vec) sell_b<12, 4>(M, in, out, m, st, out_row, bump); else sell_b<12, 1>(M, in, out, m, st, out_row, bump); break;
        case 13: sell_b<13, 1>(M, in, out, m, st, out_row, bump); break;
        case 14: sell_b<14, 2>(M, in, out, m, st, out_row, bump); break;
        case 15: sell_b<15, 3>(M, in, out, m, st, out_row, bump); break;
        case 16: if (vec) sell_b<16, 4>(M, in, out, m, st, out_row, bump); else sell_b<16, 1>(M, in, out, m, st, out_row, bump); break;
        default: throw Error(PGW_ERR_UNSUPPORTED, "block width must be in 1..16");
    }
    PGW_CHECK_LAUNCH("sparse map");
}

__global__ __launch_bounds__(256) void k_gram_finish(const uint64_t* __restrict__ part,
                                                     unsigned nblocks, uint32_t P,
                                                     uint32_t* __restrict__ out0,
                                                     uint32_t* __restrict__ out1, ModP m,
                                                     const uint32_t* __restrict__ pos,
                                                     uint32_t pos_off);

// lanes per row of k_sell_spmm_gram (the rule of sell_b at G = 1)
static int sell_gram_lanes(const SellDevice& M) {
    int lanes = 1;
    const double width = (double)M.padded / M.rows;
    while (lanes < 8 && (double)M.rows * lanes < (double)(1 << 16) && 2.0 * lanes <= width) lanes *= 2;
    return lanes;
}

unsigned sell_gram_plan(const SellDevice& M, size_t part_bytes, const ModP& m) {
    constexpr uint64_t P = 10;
    if (m.p >= 65536u || M.rows == 0) return 0;
    const uint64_t blocks = (M.rows * sell_gram_lanes(M) + 255) / 256;
    if (blocks * 2 * P * 2 * sizeof(uint64_t) > part_bytes || blocks >= (1ULL << 31)) return 0;
    return (unsigned)blocks;
}

bool launch_sell_spmm_gram(const SellDevice& M, const uint32_t* in, uint32_t* out, const uint32_t* other,
                           int mode, uint32_t* out0, uint32_t* out1, uint64_t* part, size_t part_bytes,
                           const uint32_t* pos, uint32_t pos_off, const ModP& m, hipStream_t st,
                           unsigned* nblocks) {
    constexpr uint32_t P = 10;
    auto aligned = [](const void* q) { return reinterpret_cast<uintptr_t>(q) % 16 == 0; };
    const unsigned blocks = sell_gram_plan(M, part_bytes, m);
    if (!blocks || !aligned(in) || !aligned(out) || !aligned(other) || (mode != 1 && mode != 2)) return false;
    const int lanes = sell_gram_lanes(M);
    const dim3 grid(blocks), blk(256);
#define PGW_SGF(L, FORM)                                                                                 \
    hipLaunchKernelGGL((k_sell_spmm_gram<L, kSellBatch, FORM>), grid, blk, 0, st, M.slice, M.ent, M.sbase,      \
                       M.rows, in, out, m, M.pack_bits, other, mode, part)
#define PGW_SG(L)                                                                                        \
    do {                                                                                                 \
        if (M.form == 2) PGW_SGF(L, 2); else if (M.form == 1) PGW_SGF(L, 1); else PGW_SGF(L, 0);          \
    } while (0)
    if (lanes == 8) PGW_SG(8);
    else if (lanes == 4) PGW_SG(4);
    else if (lanes == 2) PGW_SG(2);
    else PGW_SG(1);
#undef PGW_SG
#undef PGW_SGF
    if (nblocks) {
        *nblocks = (unsigned)blocks;  // the caller finishes the sums
    } else {
        hipLaunchKernelGGL(k_gram_finish, dim3(2 * P), dim3(256), 0, st, part, (unsigned)blocks, P, out0, out1,
                           m, pos, pos_off);
    }
    PGW_CHECK_LAUNCH("sparse map with projection");
    return true;
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

// v mod p for v < 2^63.  M31 (p = 2^31 - 1): 2^31 == 1, so two folds of the
// high bits onto the low 31 bits leave v' < 2^31 + 2 and one conditional
// subtraction finishes; otherwise 64-bit Barrett.
template <bool M31>
__device__ __forceinline__ uint32_t red63(uint64_t v, const ModP& m) {
    if constexpr (M31) {
        return red63_m31(v);
    } else {
        return mod_u64(v, m);
    }
}

// one round of paired 2x2 maps; forward [[1,a],[b,1+ab]], transpose [[1,b],[a,1+ab]]
// (cuda/cyclic_cuda_helpers.h:109-143).  `mid` fuses forward, Dr and transpose
// of the same round: the 2x2 block P_r^T diag(dr_f, dr_s) P_r on each pair.
// Every sum below is u*x + v*y with residues below 2^31, so below 2^63.
//
// Each lane moves CPL = 4 columns of both rows with 16-byte loads and stores
// when B is a multiple of 4 (one column otherwise), so the B / CPL lanes of a
// pair cover its two rows with one coalesced access each and a wave keeps
// 64 CPL / B pairs in flight.  (One lane per pair with four 16-byte loads per
// row ran slower: every load instruction then touches 64 scattered lines.)
template <int B>
struct MixShape {
    static constexpr int CPL = B % 4 == 0 ? 4 : 1;  // columns per lane
    static constexpr int LPP = B / CPL;             // lanes per pair
    static constexpr int PPB = 256 / LPP;           // pairs per 256-thread block
};

template <int B, int MODE, bool M31>  // MODE 0 forward, 1 transpose, 2 fused middle
__device__ __forceinline__ void rowmix_pair_cols(const pgw_rowmix_pair& pr, uint32_t* __restrict__ block,
                                                 const uint32_t* __restrict__ dr, const ModP& m, int lane) {
    constexpr int CPL = MixShape<B>::CPL;
    const uint32_t p = m.p;
    const uint32_t opab = addmod(1u, red63<M31>((uint64_t)pr.a * pr.b, m), p);
    uint32_t* fr = block + (uint64_t)pr.first * B + lane * CPL;
    uint32_t* sr = block + (uint64_t)pr.second * B + lane * CPL;
    uint32_t x[CPL], y[CPL];
    if constexpr (CPL == 4) {
        const uint4 a = *reinterpret_cast<const uint4*>(fr), c = *reinterpret_cast<const uint4*>(sr);
        x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
        y[0] = c.x; y[1] = c.y; y[2] = c.z; y[3] = c.w;
    } else {
        x[0] = *fr;
        y[0] = *sr;
    }
    uint32_t dr_f = 0, dr_s = 0;
    if constexpr (MODE == 2) {
        dr_f = dr[pr.first];
        dr_s = dr[pr.second];
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        const uint64_t xc = x[c], yc = y[c];
        uint32_t nx, ny;
        if constexpr (MODE == 1) {
            nx = red63<M31>(xc + (uint64_t)pr.b * yc, m);
            ny = red63<M31>((uint64_t)pr.a * xc + (uint64_t)opab * yc, m);
        } else {
            nx = red63<M31>(xc + (uint64_t)pr.a * yc, m);
            ny = red63<M31>((uint64_t)pr.b * xc + (uint64_t)opab * yc, m);
            if constexpr (MODE == 2) {
                const uint64_t sx = red63<M31>((uint64_t)nx * dr_f, m);
                const uint64_t sy = red63<M31>((uint64_t)ny * dr_s, m);
                nx = red63<M31>(sx + (uint64_t)pr.b * sy, m);
                ny = red63<M31>((uint64_t)pr.a * sx + (uint64_t)opab * sy, m);
            }
        }
        x[c] = nx;
        y[c] = ny;
    }
    if constexpr (CPL == 4) {
        *reinterpret_cast<uint4*>(fr) = make_uint4(x[0], x[1], x[2], x[3]);
        *reinterpret_cast<uint4*>(sr) = make_uint4(y[0], y[1], y[2], y[3]);
    } else {
        *fr = x[0];
        *sr = y[0];
    }
}

template <int B, int MODE, bool M31>
__global__ __launch_bounds__(256) void k_rowmix(const pgw_rowmix_pair* __restrict__ pairs,
                                                uint64_t count, uint32_t* __restrict__ block,
                                                const uint32_t* __restrict__ dr, ModP m) {
    constexpr int LPP = MixShape<B>::LPP, PPB = MixShape<B>::PPB;
    const uint64_t q = (uint64_t)blockIdx.x * PPB + threadIdx.x / LPP;
    if ((int)(threadIdx.x / LPP) >= PPB || q >= count) return;
    rowmix_pair_cols<B, MODE, M31>(pairs[q], block, dr, m, (int)(threadIdx.x % LPP));
}

// many jobs' row-mixing rounds in one launch; slot `count` of a job (one past
// its pairs) scales its unpaired row by Dr in the fused middle pass
template <int B, int MODE, bool M31>
__global__ __launch_bounds__(256) void k_rowmix_group(const MixJob* __restrict__ jobs,
                                                      const uint32_t* __restrict__ first,
                                                      uint32_t njobs, uint32_t round) {
    constexpr int CPL = MixShape<B>::CPL, LPP = MixShape<B>::LPP, PPB = MixShape<B>::PPB;
    const uint32_t j = group_job(first, njobs);
    const MixJob& J = jobs[j];
    const int lane = (int)(threadIdx.x % LPP);
    if ((int)(threadIdx.x / LPP) >= PPB) return;
    const uint64_t q = (uint64_t)(blockIdx.x - first[j]) * PPB + threadIdx.x / LPP;
    if (q < J.count) {
        rowmix_pair_cols<B, MODE, M31>(J.pairs[(uint64_t)round * J.count + q], J.block, J.dr, J.m, lane);
    } else if (MODE == 2 && q == J.count && J.unpaired >= 0) {
        uint32_t* x = J.block + (uint64_t)J.unpaired * B + lane * CPL;
        for (int c = 0; c < CPL; ++c) x[c] = mulmod(x[c], J.dr[J.unpaired], J.m);
    }
}

// One pass of P^T Dr P along the chains the row-mixing rounds link (rows of
// a path, or of a cycle with kMixHalo ghost copies of its rows on either
// side; see MixChains).  A workgroup gathers a window of kMixWindow chain
// positions into LDS, applies P_0, P_1, Dr, P_1^T, P_0^T to the window's
// pairs in place (the pairs of one round are disjoint, so a stage has no
// conflicts), and stores the middle kMixWindow - 2 kMixHalo positions that
// are not ghosts to their rows of `out`.  An error at a window edge (a pair
// whose partner lies outside) moves inward one position per mixing stage,
// four stages in all, so kMixHalo = 4 leaves the stored positions exact.
// Each row is read once and written once, where the three row-mixing passes
// read and write every row three times.
constexpr int kMixWindow = 256;

template <int B, bool M31>
__global__ __launch_bounds__(256) void k_mix_chain(const uint32_t* __restrict__ crow, const uint32_t* __restrict__ ca,
                                                   const uint32_t* __restrict__ cb, const uint32_t* __restrict__ cdr,
                                                   const uint8_t* __restrict__ cflag, uint64_t n,
                                                   const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                   ModP m) {
    constexpr int CPL = MixShape<B>::CPL, LPP = MixShape<B>::LPP;
    constexpr int RPP = 256 / LPP;  // rows per pass of the workgroup
    constexpr int W = kMixWindow, H = (int)kMixHalo, S = W - 2 * H;
    __shared__ uint32_t s_x[W * B];
    __shared__ uint32_t s_a[W], s_b[W], s_dr[W];
    __shared__ uint8_t s_f[W];
    const int t = threadIdx.x, rg = t / LPP, lane = t % LPP;
    const int64_t w0 = (int64_t)blockIdx.x * S - H;
    for (int i = t; i < W; i += 256) {
        const int64_t g = w0 + i;
        const bool ok = g >= 0 && (uint64_t)g < n;
        s_f[i] = ok ? cflag[g] : (uint8_t)kMixNone;
        s_a[i] = ok ? ca[g] : 0u;
        s_b[i] = ok ? cb[g] : 0u;
        s_dr[i] = ok ? cdr[g] : 0u;
        // a position at the end of the data has no link past it
        if (ok && (uint64_t)g + 1 == n) s_f[i] = (uint8_t)((s_f[i] & ~3u) | kMixNone);
    }
    if (rg < RPP) {
        for (int i = rg; i < W; i += RPP) {
            const int64_t g = w0 + i;
            uint32_t* x = s_x + i * B + lane * CPL;
            if (g >= 0 && (uint64_t)g < n) {
                const uint32_t* src = in + (uint64_t)crow[g] * B + lane * CPL;
                if constexpr (CPL == 4) {
                    *reinterpret_cast<uint4*>(x) = *reinterpret_cast<const uint4*>(src);
                } else {
                    x[0] = src[0];
                }
            } else {
#pragma unroll
                for (int c = 0; c < CPL; ++c) x[c] = 0;
            }
        }
    }
    __syncthreads();
    const uint32_t p = m.p;
    // one mixing stage: round rd, forward [[1,a],[b,1+ab]] or transpose [[1,b],[a,1+ab]] on (first, second)
    auto stage = [&](uint32_t rd, bool transpose) {
        if (rg < RPP) {
            for (int i = rg; i + 1 < W; i += RPP) {
                const uint8_t f = s_f[i];
                if ((f & 3u) != rd) continue;
                const uint32_t a = s_a[i], b = s_b[i];
                const uint32_t opab = addmod(1u, red63<M31>((uint64_t)a * b, m), p);
                uint32_t* fi = s_x + ((f & kMixFirst) ? i : i + 1) * B + lane * CPL;
                uint32_t* se = s_x + ((f & kMixFirst) ? i + 1 : i) * B + lane * CPL;
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const uint64_t xc = fi[c], yc = se[c];
                    if (transpose) {
                        fi[c] = red63<M31>(xc + (uint64_t)b * yc, m);
                        se[c] = red63<M31>((uint64_t)a * xc + (uint64_t)opab * yc, m);
                    } else {
                        fi[c] = red63<M31>(xc + (uint64_t)a * yc, m);
                        se[c] = red63<M31>((uint64_t)b * xc + (uint64_t)opab * yc, m);
                    }
                }
            }
        }
        __syncthreads();
    };
    stage(0, false);
    stage(1, false);
    if (rg < RPP)
        for (int i = rg; i < W; i += RPP) {
            uint32_t* x = s_x + i * B + lane * CPL;
#pragma unroll
            for (int c = 0; c < CPL; ++c) x[c] = red63<M31>((uint64_t)x[c] * s_dr[i], m);
        }
    __syncthreads();
    stage(1, true);
    stage(0, true);
    if (rg < RPP)
        for (int i = H + rg; i < H + S; i += RPP) {
            const int64_t g = w0 + i;
            if (g < 0 || (uint64_t)g >= n || (s_f[i] & kMixGhost)) continue;
            const uint32_t* x = s_x + i * B + lane * CPL;
            uint32_t* dst = out + (uint64_t)crow[g] * B + lane * CPL;
            if constexpr (CPL == 4) {
                *reinterpret_cast<uint4*>(dst) = *reinterpret_cast<const uint4*>(x);
            } else {
                dst[0] = x[0];
            }
        }
}

template <int B>
static void mix_chain_b(const MixChainsDevice& c, const uint32_t* in, uint32_t* out, const ModP& m, hipStream_t st) {
    constexpr int S = kMixWindow - 2 * (int)kMixHalo;
    const uint64_t blocks = (c.n + S - 1) / S;
    require(blocks < (1ULL << 31), "row-mixing chains too long for one launch");
    if (m.p == 0x7fffffffu)
        hipLaunchKernelGGL((k_mix_chain<B, true>), dim3((unsigned)blocks), dim3(256), 0, st, c.row, c.a, c.b, c.dr,
                           c.flag, c.n, in, out, m);
    else
        hipLaunchKernelGGL((k_mix_chain<B, false>), dim3((unsigned)blocks), dim3(256), 0, st, c.row, c.a, c.b, c.dr,
                           c.flag, c.n, in, out, m);
}

void launch_mix_chains(const MixChainsDevice& c, const uint32_t* in, uint32_t* out, uint32_t b, const ModP& m,
                       hipStream_t st) {
    if (c.n == 0) return;
    // 16-byte row accesses need 16-byte aligned rows
    require(b % 4 != 0 || (reinterpret_cast<uintptr_t>(in) % 16 == 0 && reinterpret_cast<uintptr_t>(out) % 16 == 0),
            "row-mixing chains need 16-byte aligned blocks");
#define PGW_CALL(BB) mix_chain_b<BB>(c, in, out, m, st)
    PGW_SWITCH_B(b, PGW_CALL)
#undef PGW_CALL
    PGW_CHECK_LAUNCH("row-mixing chains");
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
    dim3 grid(grid_for(pc, MixShape<B>::PPB)), blk(256);
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

uint64_t mix_group_blocks(const MixJob& j, uint32_t b) {
    // pairs per block (MixShape<b>::PPB); one slot past the pairs: the unpaired row
    const uint64_t ppb = 256 / (b % 4 == 0 ? b / 4 : b);
    return (j.count + 1 + ppb - 1) / ppb;
}

template <int B>
static void rowmix_group_b(const MixJob* jobs, const uint32_t* first, uint32_t njobs, uint64_t blocks,
                           int mode, bool m31, uint32_t round, hipStream_t st) {
    const dim3 grid((unsigned)blocks), blk(256);
#define PGW_RG(MODE, M31)                                                                         \
    hipLaunchKernelGGL((k_rowmix_group<B, MODE, M31>), grid, blk, 0, st, jobs, first, njobs, round)
    if (m31) {
        if (mode == 0) PGW_RG(0, true); else if (mode == 1) PGW_RG(1, true); else PGW_RG(2, true);
    } else {
        if (mode == 0) PGW_RG(0, false); else if (mode == 1) PGW_RG(1, false); else PGW_RG(2, false);
    }
#undef PGW_RG
}

void launch_rowmix_group(const MixJob* jobs, const uint32_t* first, uint32_t njobs, uint64_t blocks,
                         uint32_t b, int mode, bool m31, uint32_t round, hipStream_t st) {
    if (blocks == 0) return;
    require(blocks < (1ULL << 31), "grouped row mixing grid too large");
#define PGW_CALL(BB) rowmix_group_b<BB>(jobs, first, njobs, blocks, mode, m31, round, st)
    PGW_SWITCH_B(b, PGW_CALL)
#undef PGW_CALL
    PGW_CHECK_LAUNCH("grouped row mixing");
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
// Partial pass: each workgroup stages 64-row chunks of X and Y in LDS; its
// threads are (pair, row lane) with pair = ut(i, j) = i(2b-i+1)/2 + (j-i) and
// 256 / (b(b+1)/2) row lanes, so every thread works for any b (b = 4: 25 lanes
// of the 10 pairs).  Each keeps 64-bit sums of the low and high 32-bit words
// of its products (each product < 2^62), lanes are added in LDS, and the
// workgroup's sums land at part[(mat * P + pair) * nblocks + block].  Every
// partial and every sum of them stays below rows * 2^32 < 2^64 for rows <
// 2^32, so all of it is exact.  Finish pass: one workgroup per output entry
// adds that entry's partials (coalesced, then an LDS tree), reduces once and
// writes the packed value.

constexpr int kGramRows = 64;
constexpr unsigned kGramGrid = 1024;
constexpr unsigned kGramNoallocGrid = 256;  // workgroups of the allocation-free variant
constexpr uint64_t kGramRowsPerBlock = 256; // least rows per partial workgroup

size_t gram_scratch_bytes(uint32_t b) {
    return (size_t)kGramGrid * 2 * 256 * 2 * sizeof(uint64_t);
}

template <int B>
__device__ __forceinline__ void gram_partial_one(const uint32_t* __restrict__ x,
                                                 const uint32_t* __restrict__ y, uint64_t rows,
                                                 int mode, uint64_t* __restrict__ part, uint32_t block,
                                                 uint32_t nblocks) {
    constexpr int P = B * (B + 1) / 2;
    constexpr int L = 256 / P;  // row lanes (P <= 136)
    __shared__ uint32_t sx[kGramRows * B];
    __shared__ uint32_t sy[kGramRows * B];
    __shared__ uint64_t red[4][L * P];
    const int t = threadIdx.x;
    const bool active = t < L * P;
    const int pr = t % P, lane = t / P;
    int i = 0, rem = pr;
    while (rem >= B - i) {
        rem -= B - i;
        ++i;
    }
    const int j = i + rem;
    uint64_t lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0;
    for (uint64_t r0 = (uint64_t)block * kGramRows; r0 < rows; r0 += (uint64_t)nblocks * kGramRows) {
        const uint64_t nrow = min((uint64_t)kGramRows, rows - r0);
        __syncthreads();
        for (int e = t; e < kGramRows * B; e += 256) {
            const bool ok = (uint64_t)(e / B) < nrow;
            sx[e] = ok ? x[r0 * B + e] : 0u;
            sy[e] = ok ? y[r0 * B + e] : 0u;
        }
        __syncthreads();
        if (active) {
            for (int r = lane; r < kGramRows; r += L) {
                const uint64_t p0 = (uint64_t)sx[r * B + i] * sy[r * B + j];
                lo0 += (uint32_t)p0;
                hi0 += p0 >> 32;
                if (mode) {
                    const uint64_t p1 = mode == 1 ? (uint64_t)sy[r * B + i] * sy[r * B + j]
                                                  : (uint64_t)sx[r * B + i] * sx[r * B + j];
                    lo1 += (uint32_t)p1;
                    hi1 += p1 >> 32;
                }
            }
        }
    }
    if (active) {
        red[0][t] = lo0;
        red[1][t] = hi0;
        red[2][t] = lo1;
        red[3][t] = hi1;
    }
    __syncthreads();
    if (t < P) {
        uint64_t a[4] = {0, 0, 0, 0};
        for (int l = 0; l < L; ++l)
            for (int q = 0; q < 4; ++q) a[q] += red[q][l * P + t];
        const uint64_t nb = nblocks;
        uint64_t* o = part + ((uint64_t)t * nb + block) * 2;
        o[0] = a[0];
        o[1] = a[1];
        if (mode) {
            uint64_t* o1 = part + ((uint64_t)(P + t) * nb + block) * 2;
            o1[0] = a[2];
            o1[1] = a[3];
        }
    }
}

template <int B>
__global__ __launch_bounds__(256) void k_gram_partial(const uint32_t* __restrict__ x,
                                                      const uint32_t* __restrict__ y, uint64_t rows,
                                                      int mode, uint64_t* __restrict__ part) {
    gram_partial_one<B>(x, y, rows, mode, part, blockIdx.x, gridDim.x);
}

template <int B>
__global__ __launch_bounds__(256) void k_gram_partial_group(const GramJob* __restrict__ jobs,
                                                            const uint32_t* __restrict__ first,
                                                            uint32_t njobs, int mode) {
    const uint32_t j = group_job(first, njobs);
    const GramJob& J = jobs[j];
    gram_partial_one<B>(J.x, J.y, J.rows, mode, J.part, blockIdx.x - first[j], J.nblocks);
}

// One workgroup per output entry (mat * P + pair).  With `pos` set, out0/out1
// are offset by ut * (*pos + pos_off) entries: the sequence position is read
// on the device, so a captured graph of a whole round replays unchanged.
__device__ __forceinline__ void gram_finish_one(const uint64_t* __restrict__ part, unsigned nblocks,
                                                uint32_t P, uint32_t* __restrict__ out0,
                                                uint32_t* __restrict__ out1, const ModP& m,
                                                const uint32_t* __restrict__ pos, uint32_t pos_off,
                                                uint32_t e) {
    __shared__ uint64_t s_lo[256], s_hi[256];
    const uint32_t t = threadIdx.x;
    uint64_t lo = 0, hi = 0;
    for (uint32_t g = t; g < nblocks; g += 256) {
        const uint64_t* o = part + ((uint64_t)e * nblocks + g) * 2;
        lo += o[0];
        hi += o[1];
    }
    s_lo[t] = lo;
    s_hi[t] = hi;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (t < w) {
            s_lo[t] += s_lo[t + w];
            s_hi[t] += s_hi[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        const uint32_t mat = e / P, pr = e % P;
        uint32_t* out = mat ? out1 : out0;
        if (pos) out += (uint64_t)P * (*pos + pos_off);
        out[pr] = mod_split(s_hi[0], s_lo[0], m);
    }
}

__global__ __launch_bounds__(256) void k_gram_finish(const uint64_t* __restrict__ part,
                                                     unsigned nblocks, uint32_t P,
                                                     uint32_t* __restrict__ out0,
                                                     uint32_t* __restrict__ out1, ModP m,
                                                     const uint32_t* __restrict__ pos,
                                                     uint32_t pos_off) {
    gram_finish_one(part, nblocks, P, out0, out1, m, pos, pos_off, blockIdx.x);
}

// both halves of a sequence round at b = 4: workgroups 0-19 finish the
// first half's sums (rows *pos, *pos + 1), 20-39 the second's (*pos + 2, + 3)
__global__ __launch_bounds__(256) void k_gram_finish_round(const uint64_t* __restrict__ part_a, unsigned na,
                                                           const uint64_t* __restrict__ part_b, unsigned nb,
                                                           uint32_t* __restrict__ out0, uint32_t* __restrict__ out1,
                                                           ModP m, const uint32_t* __restrict__ pos) {
    constexpr uint32_t P = 10;
    const bool second = blockIdx.x >= 2 * P;
    gram_finish_one(second ? part_b : part_a, second ? nb : na, P, out0, out1, m, pos, second ? 2u : 0u,
                    blockIdx.x % (2 * P));
}

void launch_gram_finish_round(const uint64_t* part_a, unsigned na, const uint64_t* part_b, unsigned nb,
                              uint32_t* out0, uint32_t* out1, const ModP& m, const uint32_t* pos, hipStream_t st) {
    hipLaunchKernelGGL(k_gram_finish_round, dim3(40), dim3(256), 0, st, part_a, na, part_b, nb, out0, out1, m, pos);
    PGW_CHECK_LAUNCH("round projection finish");
}

__global__ __launch_bounds__(256) void k_gram_finish_group(const GramJob* __restrict__ jobs,
                                                           const uint32_t* __restrict__ first,
                                                           uint32_t njobs, uint32_t P,
                                                           uint32_t pos_off) {
    const uint32_t j = group_job(first, njobs);
    const GramJob& J = jobs[j];
    gram_finish_one(J.part, J.nblocks, P, J.out0, J.out1, J.m, J.pos, pos_off, blockIdx.x - first[j]);
}

__global__ void k_advance_pos_group(uint32_t* const* __restrict__ pos, uint32_t njobs, uint32_t by) {
    for (uint32_t j = threadIdx.x; j < njobs; j += blockDim.x) *pos[j] += by;
}

void launch_advance_pos_group(uint32_t* const* pos, uint32_t njobs, uint32_t by, hipStream_t st) {
    if (njobs == 0) return;
    hipLaunchKernelGGL(k_advance_pos_group, dim3(1), dim3(256), 0, st, pos, njobs, by);
    PGW_CHECK_LAUNCH("grouped advance position");
}

uint32_t gram_group_blocks(uint64_t rows) {
    return (uint32_t)std::min<uint64_t>(kGramGrid, std::max<uint64_t>(1, (rows + kGramRowsPerBlock - 1) / kGramRowsPerBlock));
}

void launch_gram_group(const GramJob* jobs, const uint32_t* first_part, const uint32_t* first_fin,
                       uint32_t njobs, uint64_t part_blocks, uint64_t fin_blocks, uint32_t b, int mode,
                       uint32_t pos_off, hipStream_t st) {
    if (njobs == 0) return;
    require(part_blocks < (1ULL << 31) && fin_blocks < (1ULL << 31), "grouped Gram grid too large");
#define PGW_CALL(BB)                                                                              \
    hipLaunchKernelGGL(k_gram_partial_group<BB>, dim3((unsigned)part_blocks), dim3(256), 0, st,   \
                       jobs, first_part, njobs, mode)
    PGW_SWITCH_B(b, PGW_CALL)
#undef PGW_CALL
    hipLaunchKernelGGL(k_gram_finish_group, dim3((unsigned)fin_blocks), dim3(256), 0, st, jobs,
                       first_fin, njobs, b * (b + 1) / 2, pos_off);
    PGW_CHECK_LAUNCH("grouped gram");
}

__global__ void k_advance_pos(uint32_t* pos, uint32_t by) {
    if (threadIdx.x == 0) *pos += by;
}

void launch_advance_pos(uint32_t* pos, uint32_t by, hipStream_t st) {
    hipLaunchKernelGGL(k_advance_pos, dim3(1), dim3(64), 0, st, pos, by);
    PGW_CHECK_LAUNCH("advance position");
}

// the two Gram kernels over at most `max_grid` workgroups (partials in `part`)
static void gram_launch(const uint32_t* x, const uint32_t* y, uint64_t rows, uint32_t b, int mode,
                        uint32_t* out0, uint32_t* out1, uint64_t* part, unsigned max_grid,
                        const ModP& m, hipStream_t st, const uint32_t* pos, uint32_t pos_off,
                        bool mfma = false) {
    // low words: rows * 2^32 < 2^64 requires rows < 2^32
    require(rows < (1ULL << 32), "Gram projection supports fewer than 2^32 rows");
    require(mode == 0 || out1 != nullptr, "Gram projection of two products needs both outputs");
    const uint32_t P = b * (b + 1) / 2;
    unsigned grid;
    if (mfma) {
        // PGW_GRAM_MAX_GRID caps the MFMA grid (more chunks per wave, so the
        // periodic folds of the accumulators run; a test knob)
        const char* env = std::getenv("PGW_GRAM_MAX_GRID");
        const long cv = env ? std::strtol(env, nullptr, 10) : 0;
        const unsigned cap = cv >= 1 ? (unsigned)std::min<long>(cv, 1L << 20) : 0u;
        grid = gram_mfma_grid(rows, b, cap ? std::min(cap, max_grid) : max_grid);
        launch_gram_mfma_partial(x, y, rows, b, mode, part, grid, m, st);
    } else {
        grid = (unsigned)std::min<uint64_t>(
            max_grid, std::max<uint64_t>(1, (rows + kGramRowsPerBlock - 1) / kGramRowsPerBlock));
        // (the grouped launch, gram_group_blocks, uses the same rule with kGramGrid)
#define PGW_CALL(BB)                                                                           \
    hipLaunchKernelGGL(k_gram_partial<BB>, dim3(grid), dim3(256), 0, st, x, y, rows, mode, part)
        PGW_SWITCH_B(b, PGW_CALL)
#undef PGW_CALL
    }
    hipLaunchKernelGGL(k_gram_finish, dim3(P * (mode ? 2 : 1)), dim3(256), 0, st, part, grid, P,
                       out0, out1, m, pos, pos_off);
    PGW_CHECK_LAUNCH("gram");
}

void launch_gram(const uint32_t* x, const uint32_t* y, uint64_t rows, uint32_t b, int mode,
                 uint32_t* out0, uint32_t* out1, uint64_t* scratch, const ModP& m,
                 hipStream_t st, const uint32_t* pos, uint32_t pos_off, bool mfma) {
    gram_launch(x, y, rows, b, mode, out0, out1, scratch, kGramGrid, m, st, pos, pos_off, mfma);
}

// ---------------------------------------------------------------------------
// Gram projection without any allocation: the partials live in a fixed
// device array (one copy per GPU, held until the process ends) instead of an
// operator's scratch.  Every partial a call reads is written by the same
// call, so nothing needs clearing; calls that overlap on one GPU share the
// array and race.
// ---------------------------------------------------------------------------

static __device__ __align__(256) uint64_t scratch[(size_t)kGramNoallocGrid * 2 * 256 * 2];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mtx;
uint64_t* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

size_t gram_noalloc_scratch_bytes() { return sizeof(scratch); }

void gram_noalloc_setup(int device) {
    require(device >= 0 && device < kMaxDevices, "device index outside the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mtx);
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;  // the array's address on the current device
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
        throw Error(PGW_ERR_HIP, std::string("hipGetSymbolAddress of the Gram scratch failed: ") +
                                     hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<uint64_t*>(ptr);
}

void launch_gram_noalloc(int device, const uint32_t* x, const uint32_t* y, uint64_t rows,
                         uint32_t b, uint32_t* out, const ModP& m, hipStream_t st) {
    uint64_t* part = nullptr;
    if (device >= 0 && device < kMaxDevices) {
        std::lock_guard<std::mutex> lk(g_scratch_mtx);
        part = g_scratch_ptr[device];
    }
    require(part != nullptr, "the Gram scratch of this device was not set up (pgw_ctx_create)");
    gram_launch(x, y, rows, b, 0, out, nullptr, part, kGramNoallocGrid, m, st, nullptr, 0);
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

// pgw_mfma.hip -- the int8-limb MFMA table contraction of the dense Koszul sweep.
//
// Residue x in [0,p) is written as a balanced base-256 number
//   x == y = sum_j d_j 256^j (mod p),  d_j in [-128, 127],  j < NL,
// with y = x if x <= MAXPOS else x - p (MAXPOS = 127*(256^NL-1)/255); every
// residue has such a representative because p <= 256^NL.  The product of two
// residues is sum_{j,l} d_j e_l 256^(j+l): the MFMA accumulates, per digit
// shift s = j + l, the int32 sum of products |d_j e_l| <= 2^14 over
// (incidences x K) terms, so an accumulator stays below 2^31 while
// incidences * K * NL < 2^17 (checked on the host).  The epilogue folds the
// shifts with w_s = 256^s mod p and reduces exactly.
//
// Digit planes of an input block: lp[sign][v][limb][c][k] (k fastest, row
// stride K_stride = round_up(K,16) bytes), sign 1 holding the digits of -x mod p
// so that the incidence sign selects a plane.  Tables: tabl[i][limb][row][k]
// with zero rows/columns up to (R_pad, K_pad), R_pad = round_up(R,16),
// K_pad = round_up(K,32).  A B fragment may read past K into the next column's
// bytes; those meet zero table entries and contribute nothing.
//
// Work decomposition: a workgroup of WR x WC wavefronts owns WR*RT row tiles
// (16 rows each) x WC*CT column tiles (16 MFMA columns = 16/B subset panels) of
// the output and walks the tables i = 0..n-1 that some panel of it is incident
// to, in K chunks of 64 (32 for the tail).  The A (table) and B (digit plane)
// tiles of a chunk are staged in LDS, double-buffered: the global loads of the
// next chunk are in flight while the MFMAs of the current one run.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <utility>
#include <vector>

#include "pgw_kernels.h"

namespace pgw {

typedef int v4i __attribute__((ext_vector_type(4)));

typedef int v4i __attribute__((ext_vector_type(4)));

template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
// f(integral_constant<int, I>) for I = 0..N-1: register-array indices stay constants
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// binomial coefficients read from a 64 x 64 device table (one per context)
struct GlobalBinom {
    const uint64_t* t;
    __device__ uint64_t operator()(int a, int b) const {
        return (a < 0 || b < 0 || b > a) ? 0ULL : t[a * 64 + b];
    }
};

struct DigitSpec {
    int64_t maxpos;
    uint64_t bias;
};

__host__ __device__ inline DigitSpec digit_spec(uint32_t nl) {
    return nl == 2 ? DigitSpec{32639, 0x8080ULL} : DigitSpec{2139062143LL, 0x80808080ULL};
}

// bytes d_j + 128 of the representative; XOR 0x80 per byte gives the int8 digit
__host__ __device__ inline uint64_t digits_biased(uint32_t x, uint32_t p, const DigitSpec& ds) {
    int64_t y = (int64_t)x <= ds.maxpos ? (int64_t)x : (int64_t)x - (int64_t)p;
    return (uint64_t)(y + (int64_t)ds.bias);
}

struct ShiftWeights {
    uint32_t w[8];
};

struct MfmaSweepArgs {
    const uint64_t* binom;  // C(a,b) table, 64 x 64
    const int8_t* tabl;
    const int8_t* lp;
    const int8_t* zero;     // >= 16 zero bytes (DMA loads of absent lines)
    uint32_t n, deg_out;
    uint64_t u_out, v_in;
    uint32_t r_out, r_pad, k_in, k_pad, k_stride;
    uint32_t row_slices;
    uint64_t col_groups;
    uint32_t xcd_affinity;  // 1: grid is 8 * ceil(col_groups / (8 / row_slices)) blocks
    uint32_t* out;
    const uint32_t* scale;
    ModP m;
    ShiftWeights sw;
};

constexpr int kLdsStride = 64;  // bytes per staged row (one 64-wide K chunk)

// Byte offset of (row, k) in a staged tile.  The four 16-byte pieces of a row
// are rotated by row/4, so 16 consecutive rows read at one k land in 16
// distinct 16-byte bank groups, and 4 consecutive whole rows written by 16
// lanes cover the 64 banks once: both directions are conflict-free.
__device__ __forceinline__ int lds_off(int row, int k) {
    return row * kLdsStride + ((((k >> 4) + (row >> 2)) & 3) << 4) + (k & 15);
}
constexpr int kMaxN = 32;

// Workgroup -> (column group, row slice), XCD-aware: hardware block ids are
// dealt round-robin to the 8 XCDs.  When the row slices divide 8, XCD x works
// only on row slice x % R, so its L2 holds just that slice of the tables; its
// column groups are spread over the 8/R XCDs sharing the slice.  The grid is
// padded to a multiple of 8; returns false for the surplus blocks, which
// leave before any barrier.
__device__ __forceinline__ bool mfma_block_tile(const MfmaSweepArgs& a, uint64_t& cg, uint32_t& rs) {
    if (a.xcd_affinity) {
        const uint32_t xcd = blockIdx.x % 8, within = blockIdx.x / 8, G = 8 / a.row_slices;
        rs = xcd % a.row_slices;
        cg = (uint64_t)within * G + xcd / a.row_slices;
        return cg < a.col_groups;
    }
    cg = blockIdx.x / a.row_slices;
    rs = blockIdx.x % a.row_slices;
    return true;
}

// Incidence codes of the workgroup's output panels: s_code[g*kMaxN + i] =
// rank of the neighbour panel | sign << 31, or ~0 when panel g is not incident
// to table i; s_any[i] = some panel is.  Ends with a barrier.
template <bool FWD, int SLOTS, int NT>
__device__ __forceinline__ void mfma_incidences(const MfmaSweepArgs& a, uint64_t cg, uint64_t* s_mask,
                                                uint32_t* s_code, uint32_t* s_any) {
    const int tid = threadIdx.x;
    GlobalBinom C{a.binom};
    if (tid < SLOTS) {
        const uint64_t u = cg * SLOTS + tid;
        s_mask[tid] = u < a.u_out ? unrank_mask(u, (int)a.n, (int)a.deg_out, C) : 0ULL;
    }
    if (tid < kMaxN) s_any[tid] = 0;
    __syncthreads();
    for (int idx = tid; idx < SLOTS * (int)a.n; idx += NT) {
        const int g = idx / (int)a.n, i = idx % (int)a.n;
        const uint64_t u = cg * SLOTS + g;
        uint32_t code = 0xffffffffu;
        if (u < a.u_out) {
            const uint64_t mask = s_mask[g];
            const bool bit = (mask >> i) & 1ULL;
            if (FWD ? !bit : bit) {
                const uint32_t neg = (uint32_t)__popcll(mask & ((1ULL << i) - 1)) & 1u;
                const uint64_t nb = FWD ? rank_mask(mask | (1ULL << i), (int)a.n, (int)a.deg_out + 1, C)
                                        : rank_mask(mask & ~(1ULL << i), (int)a.n, (int)a.deg_out - 1, C);
                code = (uint32_t)nb | (neg << 31);
                s_any[i] = 1;  // benign race: every writer stores 1
            }
        }
        s_code[g * kMaxN + i] = code;
    }
    __syncthreads();
}

// MFMAs of one staged chunk: sa = table tile (NL x ROWS rows of 64 bytes),
// sb = digit-plane tile (TILES_C x NL x 16 rows), both in the lds_off layout.
template <int B, int NL, int RT, int CT, int WR, int WC, bool WIDE>
__device__ __forceinline__ void mfma_chunk(const int8_t* sa, const int8_t* sb, const uint32_t* s_code,
                                           int i, const bool* row_on, v4i (&acc)[RT][CT][2 * NL - 1]) {
    constexpr int SPT = 16 / B;
    constexpr int ROWS = WR * RT * 16;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int wr = wave / WC, wc = wave % WC;
    const int cc = lane % 16;
    bool on[CT];
    static_for<CT>([&](auto ct) {
        bool o = false;
        static_for<SPT>([&](auto q) {
            o |= s_code[((wc * CT + ct) * SPT + q) * kMaxN + i] != 0xffffffffu;
        });
        on[ct] = __builtin_amdgcn_readfirstlane((int)o) != 0;  // uniform: scalar branch
    });
    if constexpr (WIDE) {
        const int kq = (lane / 16) * 16;
        v4i fa[RT][NL], fb[CT][NL];
        static_for<RT>([&](auto rt) {
            static_for<NL>([&](auto j) {
                fa[rt][j] = *reinterpret_cast<const v4i*>(
                    sa + (j * ROWS + wr * RT * 16 + rt * 16) * kLdsStride + lds_off(cc, kq));
            });
        });
        static_for<CT>([&](auto ct) {
            static_for<NL>([&](auto l) {
                fb[ct][l] = *reinterpret_cast<const v4i*>(
                    sb + ((wc * CT + ct) * NL + l) * 16 * kLdsStride + lds_off(cc, kq));
            });
        });
        static_for<CT>([&](auto ct) {
            if (!on[ct]) return;
            static_for<RT>([&](auto rt) {
                if (!row_on[rt]) return;
                static_for<NL>([&](auto j) {
                    static_for<NL>([&](auto l) {
                        acc[rt][ct][j + l] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                            fa[rt][j], fb[ct][l], acc[rt][ct][j + l], 0, 0, 0);
                    });
                });
            });
        });
    } else {
        const int kq = (lane / 16) * 8;
        long fa[RT][NL], fb[CT][NL];
        static_for<RT>([&](auto rt) {
            static_for<NL>([&](auto j) {
                fa[rt][j] = *reinterpret_cast<const long*>(
                    sa + (j * ROWS + wr * RT * 16 + rt * 16) * kLdsStride + lds_off(cc, kq));
            });
        });
        static_for<CT>([&](auto ct) {
            static_for<NL>([&](auto l) {
                fb[ct][l] = *reinterpret_cast<const long*>(
                    sb + ((wc * CT + ct) * NL + l) * 16 * kLdsStride + lds_off(cc, kq));
            });
        });
        static_for<CT>([&](auto ct) {
            if (!on[ct]) return;
            static_for<RT>([&](auto rt) {
                if (!row_on[rt]) return;
                static_for<NL>([&](auto j) {
                    static_for<NL>([&](auto l) {
                        acc[rt][ct][j + l] = __builtin_amdgcn_mfma_i32_16x16x32_i8(
                            fa[rt][j], fb[ct][l], acc[rt][ct][j + l], 0, 0, 0);
                    });
                });
            });
        });
    }
}

// LDS fragment reads written as one asm block per group, each ending in its
// own s_waitcnt lgkmcnt(0): the compiler then sees no LDS load that could
// alias the DMA writes into the other stages and inserts no vmcnt wait of its
// own, so the loads of the next chunks stay in flight.  Offsets are
// immediates (bytes from the per-lane base).
__device__ __forceinline__ uint32_t lds_addr(const int8_t* p) {
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) int8_t*)p;
}
template <typename T> struct DsRead;
template <> struct DsRead<v4i> { static constexpr const char* op = "ds_read_b128"; };
template <> struct DsRead<long> { static constexpr const char* op = "ds_read_b64"; };

template <int O0, int O1, int O2, int O3>
__device__ __forceinline__ void lds_read4(v4i& r0, v4i& r1, v4i& r2, v4i& r3, uint32_t base) {
    asm volatile(
        "ds_read_b128 %0, %4 offset:%5\n"
        "ds_read_b128 %1, %4 offset:%6\n"
        "ds_read_b128 %2, %4 offset:%7\n"
        "ds_read_b128 %3, %4 offset:%8\n"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(r0), "=&v"(r1), "=&v"(r2), "=&v"(r3)
        : "v"(base), "n"(O0), "n"(O1), "n"(O2), "n"(O3)
        : "memory");
}
template <int O0, int O1, int O2, int O3>
__device__ __forceinline__ void lds_read4(long& r0, long& r1, long& r2, long& r3, uint32_t base) {
    asm volatile(
        "ds_read_b64 %0, %4 offset:%5\n"
        "ds_read_b64 %1, %4 offset:%6\n"
        "ds_read_b64 %2, %4 offset:%7\n"
        "ds_read_b64 %3, %4 offset:%8\n"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(r0), "=&v"(r1), "=&v"(r2), "=&v"(r3)
        : "v"(base), "n"(O0), "n"(O1), "n"(O2), "n"(O3)
        : "memory");
}
template <int O0, int O1>
__device__ __forceinline__ void lds_read2(v4i& r0, v4i& r1, uint32_t base) {
    asm volatile(
        "ds_read_b128 %0, %2 offset:%3\n"
        "ds_read_b128 %1, %2 offset:%4\n"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(r0), "=&v"(r1)
        : "v"(base), "n"(O0), "n"(O1)
        : "memory");
}
template <int O0, int O1>
__device__ __forceinline__ void lds_read2(long& r0, long& r1, uint32_t base) {
    asm volatile(
        "ds_read_b64 %0, %2 offset:%3\n"
        "ds_read_b64 %1, %2 offset:%4\n"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(r0), "=&v"(r1)
        : "v"(base), "n"(O0), "n"(O1)
        : "memory");
}
// NL fragments at base + OFF0 + q * STEP (q < NL), NL in {2, 4}
template <int NL, int OFF0, int STEP, typename T>
__device__ __forceinline__ void lds_read_limbs(T (&f)[NL], uint32_t base) {
    if constexpr (NL == 4)
        lds_read4<OFF0, OFF0 + STEP, OFF0 + 2 * STEP, OFF0 + 3 * STEP>(f[0], f[1], f[2], f[3], base);
    else
        lds_read2<OFF0, OFF0 + STEP>(f[0], f[1], base);
}

// mfma_chunk with the fragment reads in asm (DMA-fed pipeline)
template <int B, int NL, int RT, int CT, int WR, int WC, bool WIDE>
__device__ __forceinline__ void mfma_chunk_dma(const int8_t* sa, const int8_t* sb,
                                               const uint32_t* s_code, int i, const bool* row_on,
                                               v4i (&acc)[RT][CT][2 * NL - 1]) {
    constexpr int SPT = 16 / B;
    constexpr int ROWS = WR * RT * 16;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int wr = wave / WC, wc = wave % WC;
    const int cc = lane % 16;
    bool on[CT];
    static_for<CT>([&](auto ct) {
        bool o = false;
        static_for<SPT>([&](auto q) {
            o |= s_code[((wc * CT + ct) * SPT + q) * kMaxN + i] != 0xffffffffu;
        });
        on[ct] = __builtin_amdgcn_readfirstlane((int)o) != 0;  // uniform: scalar branch
    });
    using F = typename std::conditional<WIDE, v4i, long>::type;
    const int kq = WIDE ? (lane / 16) * 16 : (lane / 16) * 8;
    const uint32_t abase = lds_addr(sa) + (uint32_t)(wr * RT * 16 * kLdsStride + lds_off(cc, kq));
    const uint32_t bbase = lds_addr(sb) + (uint32_t)(wc * CT * NL * 16 * kLdsStride + lds_off(cc, kq));
    F fa[RT][NL], fb[CT][NL];
    static_for<RT>([&](auto rt) {
        constexpr int R = decltype(rt)::value;
        lds_read_limbs<NL, R * 16 * kLdsStride, ROWS * kLdsStride>(fa[R], abase);
    });
    static_for<CT>([&](auto ct) {
        constexpr int C = decltype(ct)::value;
        lds_read_limbs<NL, C * NL * 16 * kLdsStride, 16 * kLdsStride>(fb[C], bbase);
    });
    static_for<CT>([&](auto ct) {
        if (!on[ct]) return;
        static_for<RT>([&](auto rt) {
            if (!row_on[rt]) return;
            static_for<NL>([&](auto j) {
                static_for<NL>([&](auto l) {
                    if constexpr (WIDE)
                        acc[rt][ct][j + l] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                            fa[rt][j], fb[ct][l], acc[rt][ct][j + l], 0, 0, 0);
                    else
                        acc[rt][ct][j + l] = __builtin_amdgcn_mfma_i32_16x16x32_i8(
                            fa[rt][j], fb[ct][l], acc[rt][ct][j + l], 0, 0, 0);
                });
            });
        });
    });
}

// Epilogue: lane holds D[4*(lane/16) + v][lane%16] of each 16x16 tile; fold
// the digit shifts with w_s = 256^s mod p, reduce, scale, store.
template <int B, int NL, int RT, int CT, int WR, int WC>
__device__ __forceinline__ void mfma_store(const MfmaSweepArgs& a, uint64_t cg, uint32_t wg_row0,
                                           v4i (&acc)[RT][CT][2 * NL - 1]) {
    constexpr int NS = 2 * NL - 1;
    constexpr int SPT = 16 / B;
    constexpr int SLOTS = WC * CT * SPT;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int wr = wave / WC, wc = wave % WC;
    const int cc = lane % 16;
    const int my_slot = cc / B, my_c = cc % B;
    const uint64_t off = (uint64_t)a.m.p << 32;  // multiple of p above 2^31
    static_for<RT>([&](auto rt) {
        static_for<CT>([&](auto ct) {
            const uint64_t u = cg * SLOTS + (wc * CT + ct) * SPT + my_slot;
            static_for<4>([&](auto v) {
                const uint32_t row = wg_row0 + (wr * RT + rt) * 16 + (lane / 16) * 4 + v;
                if (u >= a.u_out || row >= a.r_out) return;
                uint64_t s_lo = 0, s_hi = 0;
                static_for<NS>([&](auto s) {
                    const int32_t a32 = acc[rt][ct][s][decltype(v)::value];
                    const uint32_t r = mod_u64((uint64_t)((int64_t)a32 + (int64_t)off), a.m);
                    const uint64_t prod = (uint64_t)r * a.sw.w[s];  // < 2^62
                    if (s < 4) s_lo += prod; else s_hi += prod;     // 4 terms < 2^64
                });
                uint32_t val = addmod(mod_u64(s_lo, a.m), mod_u64(s_hi, a.m), a.m.p);
                const uint64_t orow = u * a.r_out + row;
                if (a.scale) val = mulmod(val, a.scale[orow], a.m);
                a.out[orow * B + my_c] = val;
            });
        });
    });
}

// ---------------------------------------------------------------------------
// register-staged pipeline: global -> VGPRs -> LDS, two LDS buffers
// ---------------------------------------------------------------------------
template <int B, int NL, int RT, int CT, int WR, int WC, bool FWD>
__global__ __launch_bounds__(64 * WR * WC) void k_dense_mfma(MfmaSweepArgs a) {
    constexpr int NW = WR * WC;
    constexpr int NT = 64 * NW;
    constexpr int NS = 2 * NL - 1;
    constexpr int SPT = 16 / B;                 // subset panels per column tile
    constexpr int TILES_C = WC * CT;            // column tiles per workgroup
    constexpr int SLOTS = TILES_C * SPT;        // subset panels per workgroup
    constexpr int ROWS = WR * RT * 16;          // rows per workgroup
    constexpr int A_BYTES = NL * ROWS * kLdsStride;
    constexpr int B_BYTES = TILES_C * NL * 16 * kLdsStride;
    constexpr int A_ITEMS = NL * ROWS * 4;      // 16-byte pieces per 64-wide chunk
    constexpr int B_ITEMS = TILES_C * NL * 16 * 4;
    constexpr int A_PER = (A_ITEMS + NT - 1) / NT;
    constexpr int B_PER = (B_ITEMS + NT - 1) / NT;

    __shared__ __align__(16) int8_t s_a[2][A_BYTES];
    __shared__ __align__(16) int8_t s_b[2][B_BYTES];
    __shared__ uint64_t s_mask[SLOTS];
    __shared__ uint32_t s_code[SLOTS * kMaxN];
    __shared__ uint32_t s_any[kMaxN];

    uint64_t cg;
    uint32_t rs;
    if (!mfma_block_tile(a, cg, rs)) return;
    const int tid = threadIdx.x, wave = tid / 64;
    const int wr = wave / WC;
    const uint32_t wg_row0 = rs * ROWS;
    mfma_incidences<FWD, SLOTS, NT>(a, cg, s_mask, s_code, s_any);

    v4i acc[RT][CT][NS];
    static_for<RT>([&](auto rt) {
        static_for<CT>([&](auto ct) { static_for<NS>([&](auto s) { acc[rt][ct][s] = v4i{0, 0, 0, 0}; }); });
    });

    const uint64_t plane = (uint64_t)NL * B * a.k_stride;  // bytes per panel and sign
    const int n = (int)a.n;
    const uint32_t k_pad = a.k_pad;

    // stage = (i, k0): advance to the next table that some panel here is incident to
    auto next_i = [&](int i) {
        while (i < n && !s_any[i]) ++i;
        return i;
    };
    // Per-thread constants of the staging copies, hoisted out of the stage loop:
    // the 16-byte piece of the chunk row each item moves, its offset in the
    // global table / digit plane (without the per-stage table or panel base) and
    // its LDS offset.  Thread t moves items t + q*NT; piece = item % 4.
    const int piece = tid % 4;
    uint32_t a_goff[A_PER];
    int a_soff[A_PER];
    bool a_ok[A_PER];
    static_for<A_PER>([&](auto q) {
        const int item = tid + q * NT;
        const int rest = item / 4;  // rest = j*ROWS + r
        const int r = rest % ROWS, j = rest / ROWS;
        a_ok[q] = item < A_ITEMS && wg_row0 + r < a.r_pad;
        a_goff[q] = ((uint32_t)j * a.r_pad + wg_row0 + r) * k_pad + piece * 16;
        a_soff[q] = lds_off(rest, piece * 16);
    });
    uint32_t b_goff[B_PER];
    int b_soff[B_PER], b_code_idx[B_PER];
    bool b_ok[B_PER];
    static_for<B_PER>([&](auto q) {
        const int item = tid + q * NT;
        const int rest = item / 4;  // rest = (tile*NL + l)*16 + cc
        const int c16 = rest % 16, rest2 = rest / 16;
        const int l = rest2 % NL, tile = rest2 / NL;
        b_ok[q] = item < B_ITEMS;
        b_goff[q] = ((uint32_t)l * B + c16 % B) * a.k_stride + piece * 16;
        b_soff[q] = lds_off(rest, piece * 16);
        b_code_idx[q] = b_ok[q] ? (tile * SPT + c16 / B) * kMaxN : 0;
    });
    const uint64_t table_stride = (uint64_t)NL * a.r_pad * k_pad;  // bytes per table i

    v4i ra[A_PER], rb[B_PER];
    auto load_regs = [&](int i, uint32_t k0) {
        const bool piece_ok = (k_pad - k0) >= 64 || piece < 2;  // tail chunks are 32 wide
        const int8_t* abase = a.tabl + (uint64_t)i * table_stride + k0;
        static_for<A_PER>([&](auto q) {
            v4i v = v4i{0, 0, 0, 0};
            if (a_ok[q] && piece_ok) v = *reinterpret_cast<const v4i*>(abase + a_goff[q]);
            ra[q] = v;
        });
        static_for<B_PER>([&](auto q) {
            v4i v = v4i{0, 0, 0, 0};
            const uint32_t code = s_code[b_code_idx[q] + i];
            if (b_ok[q] && piece_ok && code != 0xffffffffu)
                v = *reinterpret_cast<const v4i*>(
                    a.lp + ((uint64_t)(code >> 31) * a.v_in + (code & 0x7fffffffu)) * plane + k0 +
                    b_goff[q]);
            rb[q] = v;
        });
    };
    auto store_lds = [&](int buf) {
        static_for<A_PER>([&](auto q) {
            if (tid + q * NT < A_ITEMS) *reinterpret_cast<v4i*>(&s_a[buf][a_soff[q]]) = ra[q];
        });
        static_for<B_PER>([&](auto q) {
            if (b_ok[q]) *reinterpret_cast<v4i*>(&s_b[buf][b_soff[q]]) = rb[q];
        });
    };

    // row tiles of this wave that hold table rows (wave-uniform)
    bool row_on[RT];
    static_for<RT>([&](auto rt) { row_on[rt] = wg_row0 + (wr * RT + rt) * 16 < a.r_pad; });

    // Two phases, each a double-buffered pipeline over its stages: all 64-wide
    // chunks of every incident table, then the 32-wide tail chunks.  Keeping one
    // MFMA shape per loop keeps the accumulators in place across iterations.
    const uint32_t k_full = k_pad / 64 * 64;
    auto phase = [&](auto wide_c) {
        constexpr bool WIDE = decltype(wide_c)::value;
        const uint32_t kbeg = WIDE ? 0u : k_full, kend = WIDE ? k_full : k_pad;
        if (kbeg >= kend) return;
        int i = next_i(0);
        uint32_t k0 = kbeg;
        int buf = 0;
        if (i < n) {
            load_regs(i, k0);
            store_lds(0);
        }
        __syncthreads();
        while (i < n) {
            int ni = i;
            uint32_t nk = k0 + 64;
            if (nk >= kend) {
                ni = next_i(i + 1);
                nk = kbeg;
            }
            const bool more = ni < n;
            if (more) load_regs(ni, nk);  // in flight during the MFMAs below
            mfma_chunk<B, NL, RT, CT, WR, WC, WIDE>(s_a[buf], s_b[buf], s_code, i, row_on, acc);
            if (more) store_lds(buf ^ 1);
            __syncthreads();
            buf ^= 1;
            i = ni;
            k0 = nk;
        }
    };
    phase(std::true_type{});
    phase(std::false_type{});
    mfma_store<B, NL, RT, CT, WR, WC>(a, cg, wg_row0, acc);
}

// s_waitcnt immediate that waits for vmcnt <= N only (gfx9 encoding:
// vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] | vmcnt_hi[15:14])
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

#define PGW_LDS(p) ((__attribute__((address_space(3))) void*)(p))

// ---------------------------------------------------------------------------
// DMA-fed pipeline: global_load_lds_dwordx4 writes each lane's 16 bytes
// straight to LDS (lane L -> base + 16 L), three LDS stages, two chunks in
// flight while the MFMAs of the third run.  Each wave issues LPW loads per
// chunk (a fixed count: absent tiles, rows and incidences load a zero line),
// so s_waitcnt vmcnt(LPW) before the barrier means "my part of the chunk about
// to be used has landed".  Plain s_barrier (no fence) keeps the other chunk's
// loads in flight across it.  The stage rotation is unrolled so that every LDS
// access names its stage array statically.
// ---------------------------------------------------------------------------
template <int B, int NL, int RT, int CT, int WR, int WC, bool FWD>
__global__ __launch_bounds__(64 * WR * WC) void k_dense_mfma_dma(MfmaSweepArgs a) {
    constexpr int NW = WR * WC;
    constexpr int NT = 64 * NW;
    constexpr int NS = 2 * NL - 1;
    constexpr int SPT = 16 / B;
    constexpr int TILES_C = WC * CT;
    constexpr int SLOTS = TILES_C * SPT;
    constexpr int ROWS = WR * RT * 16;
    constexpr int A_BYTES = NL * ROWS * kLdsStride;
    constexpr int B_BYTES = TILES_C * NL * 16 * kLdsStride;
    constexpr int STAGE = A_BYTES + B_BYTES;
    constexpr int A_J = A_BYTES / 1024;  // 1 KB (one wave-wide load) per instruction
    constexpr int TOT_J = STAGE / 1024;
    static_assert(A_BYTES % 1024 == 0 && B_BYTES % 1024 == 0, "stage tiles are whole 1 KB lines");
    static_assert(TOT_J % NW == 0, "every wave issues the same number of loads per chunk");
    constexpr int LPW = TOT_J / NW;

    __shared__ __align__(16) int8_t s_st0[STAGE];
    __shared__ __align__(16) int8_t s_st1[STAGE];
    __shared__ __align__(16) int8_t s_st2[STAGE];
    __shared__ uint64_t s_mask[SLOTS];
    __shared__ uint32_t s_code[SLOTS * kMaxN];
    __shared__ uint32_t s_any[kMaxN];

    uint64_t cg;
    uint32_t rs;
    if (!mfma_block_tile(a, cg, rs)) return;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int wr = wave / WC;
    const uint32_t wg_row0 = rs * ROWS;
    mfma_incidences<FWD, SLOTS, NT>(a, cg, s_mask, s_code, s_any);

    v4i acc[RT][CT][NS];
    static_for<RT>([&](auto rt) {
        static_for<CT>([&](auto ct) { static_for<NS>([&](auto s) { acc[rt][ct][s] = v4i{0, 0, 0, 0}; }); });
    });

    const uint64_t plane = (uint64_t)NL * B * a.k_stride;
    const uint64_t table_stride = (uint64_t)NL * a.r_pad * a.k_pad;
    const int n = (int)a.n;

    // per-lane source offsets of this wave's LPW loads (instruction j = wave + NW q);
    // the LDS position of lane L is 16 L, so the lane fetches the global piece
    // that the lds_off swizzle places there
    uint32_t d_off[LPW];
    bool d_ok[LPW];
    int d_code[LPW];
    static_for<LPW>([&](auto q) {
        const int j = wave + NW * q;
        const int pos = lane & 3;
        if (j < A_J) {
            const int rest = (j * 64 + lane) >> 2;  // limb * ROWS + r
            const int pc = (pos - (rest >> 2)) & 3;
            const int r = rest % ROWS, limb = rest / ROWS;
            d_ok[q] = wg_row0 + r < a.r_pad;
            d_off[q] = ((uint32_t)limb * a.r_pad + wg_row0 + r) * a.k_pad + pc * 16;
            d_code[q] = 0;
        } else {
            const int rest = ((j - A_J) * 64 + lane) >> 2;  // (tile * NL + l) * 16 + c16
            const int pc = (pos - (rest >> 2)) & 3;
            const int c16 = rest % 16, rest2 = rest / 16;
            const int l = rest2 % NL, tile = rest2 / NL;
            d_ok[q] = true;
            d_off[q] = ((uint32_t)l * B + c16 % B) * a.k_stride + pc * 16;
            d_code[q] = (tile * SPT + c16 / B) * kMaxN;
        }
    });
    auto issue = [&](int8_t* stage, int i, uint32_t k0, bool real) {
        static_for<LPW>([&](auto q) {
            const int j = wave + NW * q;
            const int8_t* src = a.zero;
            if (real) {
                if (j < A_J) {
                    if (d_ok[q]) src = a.tabl + (uint64_t)i * table_stride + k0 + d_off[q];
                } else {
                    const uint32_t code = s_code[d_code[q] + i];
                    if (code != 0xffffffffu)
                        src = a.lp + ((uint64_t)(code >> 31) * a.v_in + (code & 0x7fffffffu)) * plane +
                              k0 + d_off[q];
                }
            }
            __builtin_amdgcn_global_load_lds(src, PGW_LDS(stage + j * 1024), 16, 0, 0);
        });
    };

    bool row_on[RT];
    static_for<RT>([&](auto rt) { row_on[rt] = wg_row0 + (wr * RT + rt) * 16 < a.r_pad; });

    auto next_i = [&](int i) {
        while (i < n && !s_any[i]) ++i;
        return i;
    };
    const uint32_t k_full = a.k_pad / 64 * 64;
    auto phase = [&](auto wide_c) {
        constexpr bool WIDE = decltype(wide_c)::value;
        const uint32_t kbeg = WIDE ? 0u : k_full, kend = WIDE ? k_full : a.k_pad;
        if (kbeg >= kend) return;
        auto advance = [&](int& i, uint32_t& k) {
            k += 64;
            if (k >= kend) {
                i = next_i(i + 1);
                k = kbeg;
            }
        };
        int ci = next_i(0);  // chunk computed next
        uint32_t ck = kbeg;
        int ii = ci;         // chunk loaded next
        uint32_t ik = kbeg;
        issue(s_st0, ii, ik, ii < n);
        if (ii < n) advance(ii, ik);
        issue(s_st1, ii, ik, ii < n);
        if (ii < n) advance(ii, ik);
        auto step = [&](int8_t* cur, int8_t* load) {
            wait_vmcnt<LPW>();
            __builtin_amdgcn_s_barrier();
            issue(load, ii, ik, ii < n);
            if (ii < n) advance(ii, ik);
            mfma_chunk_dma<B, NL, RT, CT, WR, WC, WIDE>(cur, cur + A_BYTES, s_code, ci, row_on, acc);
            advance(ci, ck);
        };
        for (;;) {
            if (ci >= n) break;
            step(s_st0, s_st2);
            if (ci >= n) break;
            step(s_st1, s_st0);
            if (ci >= n) break;
            step(s_st2, s_st1);
        }
        // the last two issues loaded zero lines: let them land before the
        // stages are reused or the workgroup ends
        wait_vmcnt<0>();
        __builtin_amdgcn_s_barrier();
    };
    phase(std::true_type{});
    phase(std::false_type{});
    mfma_store<B, NL, RT, CT, WR, WC>(a, cg, wg_row0, acc);
}

#undef PGW_LDS

template <int B, int NL, int RT, int CT, int WR, int WC, bool DMA = false>
static void dense_mfma_cfg(const DenseSweep& s, const MfmaTable& t, const int8_t* lp, uint32_t* out,
                           const uint32_t* scale, const ModP& m, hipStream_t st) {
    constexpr int SLOTS = WC * CT * (16 / B);
    constexpr int ROWS = WR * RT * 16;
    require(s.n <= (uint32_t)kMaxN, "MFMA sweep supports n <= 32");
    MfmaSweepArgs a;
    a.binom = s.binom;
    require(a.binom != nullptr, "MFMA sweep needs the device binomial table");
    a.tabl = t.d;
    a.lp = lp;
    a.zero = (const int8_t*)s.zero;
    require(!DMA || a.zero != nullptr, "DMA-fed MFMA sweep needs the zero line");
    a.n = s.n;
    a.deg_out = s.fwd ? s.k - 1 : s.k;
    a.u_out = s.u_out;
    a.v_in = s.v_in;
    a.r_out = s.r_out;
    a.r_pad = t.rows_pad;
    a.k_in = s.k_in;
    a.k_pad = t.k_pad;
    a.k_stride = limb_stride(s.k_in);
    a.row_slices = (t.rows_pad + ROWS - 1) / ROWS;
    a.col_groups = (s.u_out + SLOTS - 1) / SLOTS;
    a.out = out;
    a.scale = scale;
    a.m = m;
    uint64_t w = 1;
    for (int q = 0; q < 8; ++q) {
        a.sw.w[q] = (uint32_t)w;
        w = (w * 256) % m.p;
    }
    a.xcd_affinity = (8 % a.row_slices == 0) ? 1u : 0u;
    const uint64_t G = 8 / a.row_slices;
    const uint64_t blocks = a.xcd_affinity ? 8 * ((a.col_groups + G - 1) / G)
                                           : a.col_groups * a.row_slices;
    require(blocks < (1ULL << 31), "MFMA sweep grid too large");
#define PGW_MF(K, F)                                                                              \
    hipLaunchKernelGGL((K<B, NL, RT, CT, WR, WC, F>), dim3((unsigned)blocks), dim3(64 * WR * WC), 0, \
                       st, a)
    if constexpr (DMA) {
        if (s.fwd) PGW_MF(k_dense_mfma_dma, true); else PGW_MF(k_dense_mfma_dma, false);
    } else {
        if (s.fwd) PGW_MF(k_dense_mfma, true); else PGW_MF(k_dense_mfma, false);
    }
#undef PGW_MF
    PGW_CHECK_LAUNCH("MFMA sweep");
}

template <int B, int NL>
static void dense_mfma_pick(int cfg, const DenseSweep& s, const MfmaTable& t, const int8_t* lp,
                            uint32_t* out, const uint32_t* scale, const ModP& m, hipStream_t st) {
    // 0-2: register-staged, two LDS buffers; 3-5: DMA-fed, three LDS stages
    if constexpr (NL == 4) {
        switch (cfg) {
            case 0: dense_mfma_cfg<B, NL, 2, 2, 2, 2>(s, t, lp, out, scale, m, st); return;
            case 1: dense_mfma_cfg<B, NL, 2, 2, 4, 2>(s, t, lp, out, scale, m, st); return;
            case 2: dense_mfma_cfg<B, NL, 2, 2, 2, 4>(s, t, lp, out, scale, m, st); return;
            case 3: dense_mfma_cfg<B, NL, 2, 2, 2, 2, true>(s, t, lp, out, scale, m, st); return;
            case 4: dense_mfma_cfg<B, NL, 2, 2, 4, 2, true>(s, t, lp, out, scale, m, st); return;
            default: dense_mfma_cfg<B, NL, 2, 2, 2, 4, true>(s, t, lp, out, scale, m, st); return;
        }
    } else {
        switch (cfg) {
            case 0: dense_mfma_cfg<B, NL, 2, 4, 2, 2>(s, t, lp, out, scale, m, st); return;
            case 1: dense_mfma_cfg<B, NL, 2, 4, 4, 2>(s, t, lp, out, scale, m, st); return;
            case 2: dense_mfma_cfg<B, NL, 4, 2, 2, 4>(s, t, lp, out, scale, m, st); return;
            case 3: dense_mfma_cfg<B, NL, 2, 4, 2, 2, true>(s, t, lp, out, scale, m, st); return;
            case 4: dense_mfma_cfg<B, NL, 2, 4, 4, 2, true>(s, t, lp, out, scale, m, st); return;
            default: dense_mfma_cfg<B, NL, 4, 2, 2, 4, true>(s, t, lp, out, scale, m, st); return;
        }
    }
}

}  // namespace pgw

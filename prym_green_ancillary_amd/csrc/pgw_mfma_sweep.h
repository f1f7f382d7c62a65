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
// Digit planes of an input block: lp[sign][v][limb] blocks of B columns x
// K_pad bytes, sign 1 holding the digits of -x mod p so that the incidence
// sign selects a plane.  Tables: tabl[i][limb] blocks of R_pad rows x K_pad
// bytes with zero rows/columns up to R_pad = round_up(R,16), K_pad =
// round_up(K,32).  Both are chunk-major: 64-byte K chunks, in each chunk the
// rows (columns) follow each other at 64 bytes, then the 32-byte tail chunk
// at 32 bytes per row, so every 16-row slice of a chunk is one aligned 1 KB
// block and the staging loads are fully coalesced.  Digits past K are zero in
// the tables, so whatever the planes hold there contributes nothing.
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
#include <utility>
#include <vector>

#include "pgw_fold.h"
#include "pgw_kernels.h"

namespace pgw {

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
// dealt round-robin to the 8 XCDs, so XCD x runs blocks x, x + 8, x + 16, ...
// The col_groups x row_slices work items, in row-slice-major order, are cut
// into 8 contiguous ranges of xcd_items, one per XCD: an XCD meets one or two
// row slices, so its L2 holds just that part of the tables, whatever the
// number of slices, and it walks neighbouring output panels in dispatch
// order.  The grid is 8 * xcd_items blocks; returns false for the surplus
// blocks, which leave before any barrier.
__device__ __forceinline__ bool mfma_block_tile(const MfmaSweepArgs& a, uint32_t block, uint64_t& cg,
                                                uint32_t& rs) {
    const uint64_t item = (uint64_t)(block % 8) * a.xcd_items + block / 8;
    if (item >= a.col_groups * a.row_slices) return false;
    rs = (uint32_t)(item / a.col_groups);
    cg = item % a.col_groups;
    return true;
}

// The output's row tiles (16 rows each, r_pad / 16 of them) are spread
// evenly over the row slices: slice rs starts at row tile
// rs * base + min(rs, extra) and holds base tiles, one more for the first
// `extra` slices (base = tiles / slices, extra = tiles % slices).  A
// workgroup costs about the same staging whatever its row count, so a short
// last slice (c4's transpose: 400 rows = 128 + 128 + 128 + 16) would pay
// nearly a full slice for a few rows.  base + 1 never exceeds a workgroup's
// row tiles, since slices = ceil(tiles / (row tiles per workgroup)).
__device__ __forceinline__ void mfma_slice_rows(const MfmaSweepArgs& a, uint32_t rs, uint32_t& row0,
                                                uint32_t& tiles) {
    const uint32_t all = a.r_pad / 16, base = all / a.row_slices, extra = all % a.row_slices;
    row0 = 16 * (rs * base + (rs < extra ? rs : extra));
    tiles = base + (rs < extra ? 1u : 0u);
}

// Incidence codes of the workgroup's output panels, read from the sweep's
// incidence table (n + 1 words per output panel u, built once per operator:
// word 0 the mask of tables incident to u, word 1 + i the code of table i):
// s_code[g*kMaxN + i] = rank of the neighbour panel | sign << 31, or ~0 when
// panel g is not incident to table i or lies past the last panel;
// s_masks[0] = tables some panel meets (bit i), s_masks[1 + t] = tables the
// panels of column tile t meet.  One coalesced pass and one barrier.
template <int SLOTS, int SPT, int NT>
__device__ __forceinline__ void mfma_incidences(const MfmaSweepArgs& a, uint64_t cg, uint32_t* s_code,
                                                uint32_t* s_masks) {
    constexpr int TILES_C = SLOTS / SPT;
    const int tid = threadIdx.x, n = (int)a.n;
    const uint64_t w = (uint64_t)n + 1, u0 = cg * SLOTS;
    for (int idx = tid; idx < SLOTS * n; idx += NT) {
        const int g = idx / n, i = idx % n;
        const uint64_t u = u0 + g;
        s_code[g * kMaxN + i] = u < a.u_out ? a.inc[u * w + 1 + i] : 0xffffffffu;
    }
    if (tid <= TILES_C) {
        const int g0 = tid == 0 ? 0 : (tid - 1) * SPT, g1 = tid == 0 ? SLOTS : tid * SPT;
        uint32_t m = 0;
        for (int g = g0; g < g1; ++g)
            if (u0 + g < a.u_out) m |= a.inc[(u0 + g) * w];
        s_masks[tid] = m;
    }
    __syncthreads();
}

// Epilogue: lane holds D[4*(lane/16) + v][lane%16] of each 16x16 tile; fold
// the digit shifts to one residue (fold_digit_shifts), scale, store.
// The lane's scale factors (1 without a scale vector), all loaded before the
// epilogue's first store: a load after a store to `out` would wait for it
// (the pointers may alias).  Loading them before the K loop instead costs 16
// VGPRs there, which makes the register-A kernels spill.
template <int B, int RT, int CT, int WR, int WC>
__device__ __forceinline__ void mfma_scales(const MfmaSweepArgs& a, uint64_t cg, uint32_t wg_row0,
                                            uint32_t wg_tiles, uint32_t (&sc)[RT][CT][4]) {
    constexpr int SPT = 16 / B;
    constexpr int SLOTS = WC * CT * SPT;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int wr = wave / WC, wc = wave % WC;
    const int my_slot = (lane % 16) / B;
    static_for<RT>([&](auto rt) {
        static_for<CT>([&](auto ct) {
            const uint64_t u = cg * SLOTS + (wc * CT + ct) * SPT + my_slot;
            static_for<4>([&](auto v) {
                const uint32_t row = wg_row0 + (wr * RT + rt) * 16 + (lane / 16) * 4 + v;
                const bool mine = (uint32_t)(wr * RT + rt) < wg_tiles;
                sc[rt][ct][v] = a.scale && mine && u < a.u_out && row < a.r_out ? a.scale[u * a.r_out + row] : 1u;
            });
        });
    });
}

template <int B, int NL, int RT, int CT, int WR, int WC>
__device__ __forceinline__ void mfma_store(const MfmaSweepArgs& a, uint64_t cg, uint32_t wg_row0,
                                           uint32_t wg_tiles, v4i (&acc)[RT][CT][2 * NL - 1],
                                           const uint32_t (&sc)[RT][CT][4]) {
    constexpr int NS = 2 * NL - 1;
    constexpr int SPT = 16 / B;
    constexpr int SLOTS = WC * CT * SPT;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int wr = wave / WC, wc = wave % WC;
    const int cc = lane % 16;
    const int my_slot = cc / B, my_c = cc % B;
    const bool mersenne = NL == 4 && a.m.p == kM31;
    static_for<RT>([&](auto rt) {
        static_for<CT>([&](auto ct) {
            const uint64_t u = cg * SLOTS + (wc * CT + ct) * SPT + my_slot;
            static_for<4>([&](auto v) {
                const uint32_t row = wg_row0 + (wr * RT + rt) * 16 + (lane / 16) * 4 + v;
                // rows of this slice only: the tiles after wg_tiles belong to the next one
                if ((uint32_t)(wr * RT + rt) >= wg_tiles || u >= a.u_out || row >= a.r_out) return;
                int32_t x[NS];
                static_for<NS>([&](auto s) { x[s] = acc[rt][ct][s][decltype(v)::value]; });
                uint32_t val = fold_digit_shifts<NL>(x, a.m, a.sw.w);
                const uint64_t orow = u * a.r_out + row;
                if (a.scale) {
                    const uint32_t f = sc[rt][ct][v];
                    val = mersenne ? red63_m31((uint64_t)val * f) : mulmod(val, f, a.m);  // product < 2^62
                }
                a.out[orow * B + my_c] = val;
            });
        });
    });
}

// ---------------------------------------------------------------------------
// register-staged pipeline: global -> VGPRs -> LDS, two LDS buffers
//
// The K loop walks 64-byte chunks: every full 64-wide chunk of each incident
// table, then the 32-wide tails (K_pad % 64 == 32) two tables at a time --
// bytes 0..31 of a tail chunk hold table i1's tail and the panels' i1
// neighbours, bytes 32..63 table i2's and the i2 neighbours, so one 64-deep
// MFMA adds both products and every chunk has the same shape.
// ---------------------------------------------------------------------------
template <int B, int NL, int RT, int CT, int WR, int WC, bool FWD>
__global__ __launch_bounds__(64 * WR * WC, 2) void k_dense_mfma(MfmaSweepArgs a) {
    const uint32_t block = blockIdx.x;
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
    __shared__ uint32_t s_code[SLOTS * kMaxN];
    __shared__ uint32_t s_masks[1 + TILES_C];

    uint64_t cg;
    uint32_t rs;
    if (!mfma_block_tile(a, block, cg, rs)) return;
    const int tid = threadIdx.x, wave = tid / 64;
    const int wr = wave / WC, wc = wave % WC;
    uint32_t wg_row0, wg_tiles;
    mfma_slice_rows(a, rs, wg_row0, wg_tiles);
    mfma_incidences<SLOTS, SPT, NT>(a, cg, s_code, s_masks);

    v4i acc[RT][CT][NS];
    static_for<RT>([&](auto rt) {
        static_for<CT>([&](auto ct) { static_for<NS>([&](auto s) { acc[rt][ct][s] = v4i{0, 0, 0, 0}; }); });
    });

    const uint32_t k_pad = a.k_pad;
    const uint64_t plane = (uint64_t)NL * B * k_pad;     // bytes per panel and sign
    const uint64_t limb_block_b = (uint64_t)B * k_pad;     // bytes per panel limb
    const uint64_t limb_block_a = (uint64_t)a.r_pad * k_pad;  // bytes per table limb
    const int n = (int)a.n;
    const uint32_t k_full = k_pad / 64 * 64;  // tail (k_pad % 64 == 32) starts here
    const bool has_tail = k_full < k_pad;

    // incidence masks in scalar registers (bit i: table i): tables some panel
    // of the workgroup meets, and those the panels of each of this wave's
    // column tiles meet
    const uint32_t any_mask = __builtin_amdgcn_readfirstlane(s_masks[0]);
    uint32_t on_mask[CT];
    static_for<CT>([&](auto ct) { on_mask[ct] = __builtin_amdgcn_readfirstlane(s_masks[1 + wc * CT + ct]); });
    auto next_i = [&](int i) {  // first incident table >= i, or n
        const uint32_t rest = i < 32 ? any_mask >> i : 0u;
        return rest ? i + __builtin_ctz(rest) : n;
    };

    // Per-thread constants of the staging copies: the 16-byte piece of the
    // chunk row each item moves (piece = item % 4), its row offset in the
    // global table / digit plane and its LDS offset.  Thread t moves items
    // t + q*NT.
    const int piece = tid % 4;
    uint64_t a_loff[A_PER];  // limb block of the item's table row
    uint32_t a_row[A_PER];
    int a_soff[A_PER];
    bool a_ok[A_PER];
    static_for<A_PER>([&](auto q) {
        const int item = tid + q * NT;
        const int rest = item / 4;  // rest = j*ROWS + r
        const int r = rest % ROWS, j = rest / ROWS;
        a_ok[q] = item < A_ITEMS && (uint32_t)r < wg_tiles * 16;
        a_loff[q] = (uint64_t)j * limb_block_a;
        a_row[q] = wg_row0 + r;
        a_soff[q] = lds_off(rest, piece * 16);
    });
    uint32_t b_loff[B_PER], b_col[B_PER];  // limb block and column of the item's panel row
    int b_soff[B_PER], b_code_idx[B_PER];
    bool b_ok[B_PER];
    static_for<B_PER>([&](auto q) {
        const int item = tid + q * NT;
        const int rest = item / 4;  // rest = (tile*NL + l)*16 + cc
        const int c16 = rest % 16, rest2 = rest / 16;
        const int l = rest2 % NL, tile = rest2 / NL;
        b_ok[q] = item < B_ITEMS;
        b_loff[q] = (uint32_t)(l * limb_block_b);
        b_col[q] = (uint32_t)(c16 % B);
        b_soff[q] = lds_off(rest, piece * 16);
        b_code_idx[q] = b_ok[q] ? (tile * SPT + c16 / B) * kMaxN : 0;
    });
    const uint64_t table_stride = (uint64_t)NL * a.r_pad * k_pad;  // bytes per table i

    // chunk (i, i2, k0): a full chunk of table i at k0 (i2 unused), or a tail
    // pair (k0 == k_full) of tables i and i2 (i2 == n: none)
    v4i ra[A_PER], rb[B_PER];
    auto load_regs = [&](int i, int i2, uint32_t k0) {
        // chunk-major addressing: row r (column c) of the chunk at k0 sits at
        // chunk_base + r * width + piece offset, width 64 (32 in the tail)
        const bool tail = k0 == k_full;
        const int t = (tail && piece >= 2) ? i2 : i;                      // this piece's table
        const uint32_t width = tail ? 32u : 64u;
        const uint32_t poff = tail ? (piece & 1) * 16 : piece * 16;
        const bool t_ok = t < n;
        const uint64_t a_chunk = (uint64_t)k0 * a.r_pad;                  // chunk base in a limb block
        const uint64_t b_chunk = (uint64_t)k0 * B;
        const int8_t* abase = a.tabl + (uint64_t)(t_ok ? t : 0) * table_stride + a_chunk + poff;
        static_for<A_PER>([&](auto q) {
            v4i v = v4i{0, 0, 0, 0};
            if (a_ok[q] && t_ok)
                v = *reinterpret_cast<const v4i*>(abase + a_loff[q] + (uint64_t)a_row[q] * width);
            ra[q] = v;
        });
        static_for<B_PER>([&](auto q) {
            v4i v = v4i{0, 0, 0, 0};
            const uint32_t code = t_ok ? s_code[b_code_idx[q] + t] : 0xffffffffu;
            if (b_ok[q] && code != 0xffffffffu)
                v = *reinterpret_cast<const v4i*>(
                    a.lp + ((uint64_t)(code >> 31) * a.v_in + (code & 0x7fffffffu)) * plane + b_loff[q] +
                    b_chunk + (uint64_t)b_col[q] * width + poff);
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
    static_for<RT>([&](auto rt) {
        row_on[rt] = __builtin_amdgcn_readfirstlane((int)((uint32_t)(wr * RT + rt) < wg_tiles)) != 0;
    });

    // chunk stream: full chunks of every incident table, then tail pairs
    auto first = [&](int& i, int& i2, uint32_t& k0) {
        i = next_i(0);
        if (k_full > 0) {
            i2 = n;
            k0 = 0;
        } else {
            i2 = i < n ? next_i(i + 1) : n;
            k0 = k_full;
        }
    };
    auto advance = [&](int& i, int& i2, uint32_t& k0) {
        if (k0 != k_full) {  // full chunk
            k0 += 64;
            if (k0 < k_full) return;
            i = next_i(i + 1);
            if (i < n) {
                k0 = 0;
                return;
            }
            if (!has_tail) return;  // i == n: done
            i = next_i(0);
            i2 = i < n ? next_i(i + 1) : n;
            k0 = k_full;
            return;
        }
        i = i2 < n ? next_i(i2 + 1) : n;  // next tail pair
        i2 = i < n ? next_i(i + 1) : n;
    };
    auto chunk_on = [&](uint32_t mask, int i, int i2, uint32_t k0) {
        const uint32_t both = (1u << i) | ((k0 == k_full && i2 < n) ? (1u << i2) : 0u);
        return (mask & both) != 0;
    };

    const int lane = tid % 64, cc = lane % 16;
    const int kq = (lane / 16) * 16;
    int i, i2;
    uint32_t k0;
    first(i, i2, k0);
    int buf = 0;
    if (i < n) {
        load_regs(i, i2, k0);
        store_lds(0);
    }
    __syncthreads();
    while (i < n) {
        int ni = i, ni2 = i2;
        uint32_t nk = k0;
        advance(ni, ni2, nk);
        const bool more = ni < n;
        if (more) load_regs(ni, ni2, nk);  // in flight during the MFMAs below

        const int8_t* sa = s_a[buf];
        const int8_t* sb = s_b[buf];
        // fragment reads in the order the MFMA groups below consume them
        // (column tile 0 with every row tile, then the other column tiles),
        // so that each group waits only for its own reads
        v4i fa[RT][NL], fb[CT][NL];
        static_for<CT>([&](auto ct) {
            static_for<NL>([&](auto l) {
                fb[ct][l] = *reinterpret_cast<const v4i*>(
                    sb + ((wc * CT + ct) * NL + l) * 16 * kLdsStride + lds_off(cc, kq));
            });
            if constexpr (decltype(ct)::value == 0) {
                static_for<RT>([&](auto rt) {
                    static_for<NL>([&](auto j) {
                        fa[rt][j] = *reinterpret_cast<const v4i*>(
                            sa + (j * ROWS + wr * RT * 16 + rt * 16) * kLdsStride + lds_off(cc, kq));
                    });
                });
            }
        });
        static_for<CT>([&](auto ct) {
            if (!chunk_on(on_mask[ct], i, i2, k0)) return;
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
        if (more) store_lds(buf ^ 1);
        __syncthreads();
        buf ^= 1;
        i = ni;
        i2 = ni2;
        k0 = nk;
    }
    uint32_t sc[RT][CT][4];
    mfma_scales<B, RT, CT, WR, WC>(a, cg, wg_row0, wg_tiles, sc);
    mfma_store<B, NL, RT, CT, WR, WC>(a, cg, wg_row0, wg_tiles, acc, sc);
}

// ---------------------------------------------------------------------------
// direct-to-LDS variant: three LDS stages fed by global_load_lds
//
// Same tiles and MFMA schedule as above, but the staging copies go straight
// from global memory into LDS (global_load_lds_dwordx4: every lane moves 16
// bytes, a wave fills one contiguous 1 KB block), so they need no VGPRs and
// no LDS store instructions, and three stages keep two chunks of loads in
// flight: at chunk c the loads of chunk c+2 are issued into the stage chunk
// c-1 used, while the MFMAs of chunk c run.  Each wave waits (vmcnt) for its
// own loads of chunk c, then the workgroup barrier makes every wave's loads
// visible.  The 16-byte piece a lane moves is picked so that the block lands
// in the swizzled layout of lds_off().  Loads that have no data (rows past
// the table, panels not incident to the table, the missing partner of a
// tail pair) read a 1 KB zero block.
// ---------------------------------------------------------------------------

// workgroup barrier that orders LDS accesses only (global loads stay in flight)
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One global_load_lds_dwordx4: lane L copies 16 bytes from src (its own
// address) to LDS byte lds + 16 L.  Written as inline asm because the
// compiler, which cannot tell which LDS buffer such a load writes, would
// otherwise wait for every one in flight before any LDS read; the kernel
// waits for them itself (wait_vmcnt before its barrier).  The instruction
// returns nothing to VGPRs, so no output register is left pending.
__device__ __forceinline__ void load_to_lds(const void* src, int8_t* lds) {
    const uint32_t m0 = __builtin_amdgcn_readfirstlane(
        (uint32_t)(uintptr_t)(__attribute__((address_space(3))) int8_t*)lds);
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(src), "s"(m0)
                 : "memory", "m0");
}

// s_waitcnt with only the vector-memory counter constrained (at most n loads
// of this wave still in flight); n is a compile-time constant below 64
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// Waves per SIMD the direct-to-LDS kernel is compiled for: two workgroups
// per CU when two copies of its LDS (stages, incidence codes and masks) fit
// in the 160 KB of a CU, so one workgroup's staging and fragment reads run
// under the other's MFMAs; at least two waves per SIMD either way.
template <int B, int NL, int RT, int CT, int WR, int WC, int STG>
constexpr int dma_waves_per_simd() {
    constexpr int rows = WR * RT * 16, tiles_c = WC * CT, slots = tiles_c * (16 / B);
    constexpr int stage = NL * rows * kLdsStride + tiles_c * NL * 16 * kLdsStride;
    constexpr int lds = STG * stage + slots * 4 * kMaxN + 4 * (1 + tiles_c) + 64;
    constexpr int wgs = 2 * lds <= 160 * 1024 ? 2 : 1;
    constexpr int w = wgs * WR * WC / 4;
    return w < 2 ? 2 : w;
}

template <int B, int NL, int RT, int CT, int WR, int WC, int STG, bool FWD>
__global__ __launch_bounds__(64 * WR * WC, (dma_waves_per_simd<B, NL, RT, CT, WR, WC, STG>()))
void k_dense_mfma_dma(MfmaSweepArgs a) {
    const uint32_t block = blockIdx.x;
    constexpr int NW = WR * WC;
    constexpr int NT = 64 * NW;
    constexpr int NS = 2 * NL - 1;
    constexpr int SPT = 16 / B;
    constexpr int TILES_C = WC * CT;
    constexpr int SLOTS = TILES_C * SPT;
    constexpr int ROWS = WR * RT * 16;
    constexpr int A_BYTES = NL * ROWS * kLdsStride;
    constexpr int B_BYTES = TILES_C * NL * 16 * kLdsStride;
    constexpr int A_BLOCKS = A_BYTES / 1024, B_BLOCKS = B_BYTES / 1024;  // 1 KB blocks per stage
    static_assert((A_BLOCKS + B_BLOCKS) % NW == 0, "blocks must split evenly over the waves");
    static_assert(STG == 2 || STG == 3, "two or three LDS stages");
    constexpr int PER = (A_BLOCKS + B_BLOCKS) / NW;  // loads per wave and chunk

    // one array per stage (not one indexed array), stage indices known at
    // compile time: the stage read and the stages being loaded stay apart
    __shared__ __align__(1024) int8_t s_st0[A_BYTES + B_BYTES];
    __shared__ __align__(1024) int8_t s_st1[A_BYTES + B_BYTES];
    __shared__ __align__(1024) int8_t s_st2[STG == 3 ? A_BYTES + B_BYTES : 16];
    __shared__ uint32_t s_code[SLOTS * kMaxN];
    __shared__ uint32_t s_masks[1 + TILES_C];

    uint64_t cg;
    uint32_t rs;
    if (!mfma_block_tile(a, block, cg, rs)) return;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int wr = wave / WC, wc = wave % WC;
    uint32_t wg_row0, wg_tiles;
    mfma_slice_rows(a, rs, wg_row0, wg_tiles);
    mfma_incidences<SLOTS, SPT, NT>(a, cg, s_code, s_masks);

    v4i acc[RT][CT][NS];
    static_for<RT>([&](auto rt) {
        static_for<CT>([&](auto ct) { static_for<NS>([&](auto s) { acc[rt][ct][s] = v4i{0, 0, 0, 0}; }); });
    });

    const uint32_t k_pad = a.k_pad;
    const uint64_t plane = (uint64_t)NL * B * k_pad;
    const uint32_t limb_block_b = B * k_pad;
    const uint32_t limb_block_a = a.r_pad * k_pad;
    const uint32_t table_stride = NL * limb_block_a;
    const int n = (int)a.n;
    const uint32_t k_full = k_pad / 64 * 64;
    const bool has_tail = k_full < k_pad;
    const uint32_t any_mask = __builtin_amdgcn_readfirstlane(s_masks[0]);
    uint32_t on_mask[CT];
    static_for<CT>([&](auto ct) { on_mask[ct] = __builtin_amdgcn_readfirstlane(s_masks[1 + wc * CT + ct]); });
    auto next_i = [&](int i) {
        const uint32_t rest = i < 32 ? any_mask >> i : 0u;
        return rest ? i + __builtin_ctz(rest) : n;
    };
    auto chunk_on = [&](uint32_t mask, int i, int i2, uint32_t k0) {
        const uint32_t both = (1u << i) | ((k0 == k_full && i2 < n) ? (1u << i2) : 0u);
        return (mask & both) != 0;
    };

    // lane L of a load fills bytes [16L, 16L+16) of a 1 KB block: staged row
    // L/4 of the block, physical piece L%4, i.e. logical piece (L%4 - row/4) & 3
    const int8_t* zero = a.zero + lane * 16;
    const uint32_t zero_off = (uint32_t)(zero - a.tabl);
    auto stage = [&](auto st) -> int8_t* {
        constexpr int S = decltype(st)::value % STG;
        if constexpr (S == 0) return s_st0;
        else if constexpr (S == 1) return s_st1;
        else return s_st2;
    };
    auto issue = [&](int8_t* dst, int i, int i2, uint32_t k0) {
        const bool tail = k0 == k_full;
        const uint32_t wshift = tail ? 5u : 6u;  // bytes per chunk row: 32 in the tail, else 64
        static_for<PER>([&](auto q) {
            constexpr int Q = decltype(q)::value;
            const int blk = wave + Q * NW;                    // wave-uniform
            const int row_s = blk * 16 + lane / 4;            // staged row of the stage
            const int lpc = ((lane % 4) - (row_s >> 2)) & 3;  // logical 16-byte piece
            const int t = (tail && lpc >= 2) ? i2 : i;
            const uint32_t poff = tail ? (lpc & 1) * 16 : lpc * 16;
            const int8_t* src;
            // blocks Q*NW .. Q*NW+NW-1: all table blocks, all panel blocks, or mixed
            constexpr bool all_a = (Q + 1) * NW <= A_BLOCKS, all_b = Q * NW >= A_BLOCKS;
            if (all_a || (!all_b && blk < A_BLOCKS)) {  // table rows: staged row = j * ROWS + r
                const int r = row_s % ROWS, j = row_s / ROWS;
                const uint32_t row = wg_row0 + r;
                const uint32_t off = (t < n && (uint32_t)r < wg_tiles * 16)
                                         ? (uint32_t)t * table_stride + k0 * a.r_pad + j * limb_block_a +
                                               (row << wshift) + poff
                                         : zero_off;
                src = a.tabl + off;
            } else {  // panel columns: staged row - A rows = (tile * NL + l) * 16 + c16
                const int rb = row_s - A_BLOCKS * 16;
                const int c16 = rb % 16, l = rb / 16 % NL, tile = rb / 16 / NL;
                const uint32_t cd = t < n ? s_code[(tile * SPT + c16 / B) * kMaxN + t] : 0xffffffffu;
                src = cd != 0xffffffffu
                          ? a.lp + ((uint64_t)(cd >> 31) * a.v_in + (cd & 0x7fffffffu)) * plane +
                                ((uint64_t)k0 * B + l * limb_block_b + ((uint32_t)(c16 % B) << wshift) + poff)
                          : zero;
            }
            load_to_lds(src, dst + blk * 1024);
        });
    };

    auto first = [&](int& i, int& i2, uint32_t& k0) {
        i = next_i(0);
        if (k_full > 0) {
            i2 = n;
            k0 = 0;
        } else {
            i2 = i < n ? next_i(i + 1) : n;
            k0 = k_full;
        }
    };
    auto advance = [&](int& i, int& i2, uint32_t& k0) {
        if (k0 != k_full) {
            k0 += 64;
            if (k0 < k_full) return;
            i = next_i(i + 1);
            if (i < n) {
                k0 = 0;
                return;
            }
            if (!has_tail) return;
            i = next_i(0);
            i2 = i < n ? next_i(i + 1) : n;
            k0 = k_full;
            return;
        }
        i = i2 < n ? next_i(i2 + 1) : n;
        i2 = i < n ? next_i(i + 1) : n;
    };

    bool row_on[RT];
    static_for<RT>([&](auto rt) {
        row_on[rt] = __builtin_amdgcn_readfirstlane((int)((uint32_t)(wr * RT + rt) < wg_tiles)) != 0;
    });

    const int cc = lane % 16, kq = (lane / 16) * 16;
    // chunk c = (i, i2, k0) in stage ST (c mod 3); c+1 = (j1, j12, q1); false after the last chunk
    int i, i2, j1, j12;
    uint32_t k0, q1;
    auto step = [&](auto ST) {
        constexpr int S = decltype(ST)::value;
        // this wave's loads of chunk c are done (with three stages chunk
        // c+1's may stay in flight); after the barrier every wave's are, and
        // the stage of chunk c-1 is free
        if (STG == 3 && j1 < n) wait_vmcnt<PER>(); else wait_vmcnt<0>();
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        int j2 = j1, j22 = j12;
        uint32_t q2 = q1;
        if constexpr (STG == 3) {
            if (j1 < n) {
                advance(j2, j22, q2);
                if (j2 < n) issue(stage(std::integral_constant<int, S + 2>{}), j2, j22, q2);
            }
        } else {
            if (j1 < n) {
                issue(stage(std::integral_constant<int, S + 1>{}), j1, j12, q1);
                advance(j2, j22, q2);
            }
        }
        const int8_t* sa = stage(ST);
        const int8_t* sb = sa + A_BYTES;
        v4i fa[RT][NL], fb[CT][NL];
        static_for<CT>([&](auto ct) {
            static_for<NL>([&](auto l) {
                fb[ct][l] = *reinterpret_cast<const v4i*>(sb + ((wc * CT + ct) * NL + l) * 16 * kLdsStride +
                                                          lds_off(cc, kq));
            });
            if constexpr (decltype(ct)::value == 0) {
                static_for<RT>([&](auto rt) {
                    static_for<NL>([&](auto j) {
                        fa[rt][j] = *reinterpret_cast<const v4i*>(
                            sa + (j * ROWS + wr * RT * 16 + rt * 16) * kLdsStride + lds_off(cc, kq));
                    });
                });
            }
        });
        static_for<CT>([&](auto ct) {
            if (!chunk_on(on_mask[ct], i, i2, k0)) return;
            static_for<RT>([&](auto rt) {
                if (!row_on[rt]) return;
                static_for<NL>([&](auto j) {
                    static_for<NL>([&](auto l) {
                        acc[rt][ct][j + l] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[rt][j], fb[ct][l],
                                                                                  acc[rt][ct][j + l], 0, 0, 0);
                    });
                });
            });
        });
        const bool more = j1 < n;
        i = j1;
        i2 = j12;
        k0 = q1;
        j1 = j2;
        j12 = j22;
        q1 = q2;
        return more;
    };
    first(i, i2, k0);
    if (i < n) {
        j1 = i;
        j12 = i2;
        q1 = k0;
        advance(j1, j12, q1);
        issue(s_st0, i, i2, k0);
        if constexpr (STG == 3) {
            if (j1 < n) issue(s_st1, j1, j12, q1);
            while (step(std::integral_constant<int, 0>{}) && step(std::integral_constant<int, 1>{}) &&
                   step(std::integral_constant<int, 2>{})) {
            }
        } else {
            while (step(std::integral_constant<int, 0>{}) && step(std::integral_constant<int, 1>{})) {
            }
        }
    }
    uint32_t sc[RT][CT][4];
    mfma_scales<B, RT, CT, WR, WC>(a, cg, wg_row0, wg_tiles, sc);
    mfma_store<B, NL, RT, CT, WR, WC>(a, cg, wg_row0, wg_tiles, acc, sc);
}

// ---------------------------------------------------------------------------
// register-A variant: table fragments straight from global memory to VGPRs
//
// A wave's table rows are its own, so staging them in LDS buys no reuse and
// costs LDS bandwidth twice (the staging write and the fragment read).  Here
// every wave loads the RT x NL table fragments of chunk c+1 with plain vector
// loads (per fragment one coalesced 1 KB block: 16 chunk rows of 64 bytes;
// in a tail chunk lanes 0-31 read table i's 32-byte rows, lanes 32-63 table
// i2's) into a second register set while the MFMAs of chunk c run.  Only the
// digit-plane tile, which all waves of the workgroup read, goes through LDS:
// direct-to-LDS loads into STG stages, as in the kernel above.  With three
// stages a wave issues, at chunk c, the fragments of c+1 and then the staging
// of c+2, so at chunk c+1 it waits for all but its last PER loads.
// ---------------------------------------------------------------------------
template <int B, int NL, int RT, int CT, int WR, int WC, int STG, bool FWD>
__device__ __forceinline__ void dense_mfma_areg_body(const MfmaSweepArgs& a, uint32_t block) {
    constexpr int NW = WR * WC;
    constexpr int NT = 64 * NW;
    constexpr int NS = 2 * NL - 1;
    constexpr int SPT = 16 / B;
    constexpr int TILES_C = WC * CT;
    constexpr int SLOTS = TILES_C * SPT;
    constexpr int B_BYTES = TILES_C * NL * 16 * kLdsStride;
    constexpr int B_BLOCKS = B_BYTES / 1024;
    constexpr int PER = (B_BLOCKS + NW - 1) / NW;  // staging loads per wave and chunk (at most)
    static_assert(STG == 2 || STG == 3, "two or three LDS stages");
    // the partial wait of three stages counts every wave's staging loads
    static_assert(STG == 2 || B_BLOCKS % NW == 0, "three stages need equal staging per wave");

    __shared__ __align__(1024) int8_t s_st0[B_BYTES];
    __shared__ __align__(1024) int8_t s_st1[B_BYTES];
    __shared__ __align__(1024) int8_t s_st2[STG == 3 ? B_BYTES : 16];
    __shared__ uint32_t s_code[SLOTS * kMaxN];
    __shared__ uint32_t s_masks[1 + TILES_C];
    __shared__ const int8_t* s_bptr[SLOTS * (kMaxN + 1)];  // digit plane of panel g's table-t neighbour

    uint64_t cg;
    uint32_t rs;
    if (!mfma_block_tile(a, block, cg, rs)) return;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int wr = wave / WC, wc = wave % WC;
    uint32_t wg_row0, wg_tiles;
    mfma_slice_rows(a, rs, wg_row0, wg_tiles);
    mfma_incidences<SLOTS, SPT, NT>(a, cg, s_code, s_masks);
    // s_bptr[g * (kMaxN + 1) + t]: base of the digit plane (sign and limb 0)
    // that panel g reads at table t, or the zero table when g is not incident
    // to t (t = n included: the missing partner of a tail pair).  The zero
    // table spans NL * R_pad * K_pad >= NL * B * K_pad bytes, one whole plane.
    {
        const uint64_t plane_b = (uint64_t)NL * B * a.k_pad;
        const int np1 = (int)a.n + 1;
        for (int idx = tid; idx < SLOTS * np1; idx += NT) {
            const int g = idx / np1, t = idx % np1;
            const uint32_t cd = t < (int)a.n ? s_code[g * kMaxN + t] : 0xffffffffu;
            s_bptr[g * (kMaxN + 1) + t] =
                cd != 0xffffffffu ? a.lp + ((uint64_t)(cd >> 31) * a.v_in + (cd & 0x7fffffffu)) * plane_b : a.zero;
        }
        __syncthreads();
    }

    v4i acc[RT][CT][NS];
    static_for<RT>([&](auto rt) {
        static_for<CT>([&](auto ct) { static_for<NS>([&](auto s) { acc[rt][ct][s] = v4i{0, 0, 0, 0}; }); });
    });

    const uint32_t k_pad = a.k_pad;
    const uint32_t limb_block_b = B * k_pad;
    const uint32_t limb_block_a = a.r_pad * k_pad;
    const uint32_t table_stride = NL * limb_block_a;
    const int n = (int)a.n;
    const uint32_t k_full = k_pad / 64 * 64;
    const bool has_tail = k_full < k_pad;
    const uint32_t any_mask = __builtin_amdgcn_readfirstlane(s_masks[0]);
    uint32_t on_mask[CT];
    static_for<CT>([&](auto ct) { on_mask[ct] = __builtin_amdgcn_readfirstlane(s_masks[1 + wc * CT + ct]); });
    auto next_i = [&](int i) {
        const uint32_t rest = i < 32 ? any_mask >> i : 0u;
        return rest ? i + __builtin_ctz(rest) : n;
    };
    auto chunk_on = [&](uint32_t mask, int i, int i2, uint32_t k0) {
        const uint32_t both = (1u << i) | ((k0 == k_full && i2 < n) ? (1u << i2) : 0u);
        return (mask & both) != 0;
    };
    bool row_on[RT];
    static_for<RT>([&](auto rt) {
        row_on[rt] = __builtin_amdgcn_readfirstlane((int)((uint32_t)(wr * RT + rt) < wg_tiles)) != 0;
    });
    const int cc = lane % 16, kq = (lane / 16) * 16;

    // Staging load q of this lane: block blk = wave + q * NW of the stage
    // holds staged rows (tile * NL + l) * 16 + c16; the lane fills logical
    // 16-byte piece lpc of row c16, i.e. panel g = tile * SPT + c16 / B, limb
    // l, column c16 % B, chunk bytes 16 lpc (a tail chunk: table i's 32 bytes
    // in pieces 0-1, table i2's in pieces 2-3).  Offsets within the plane:
    int b_g[PER];
    bool b_hi[PER];                         // tail: piece of table i2
    uint32_t b_full[PER], b_tail[PER];      // plane offset without k0 * B, full / tail chunk
    static_for<PER>([&](auto q) {
        const int blk = wave + (int)decltype(q)::value * NW;
        const int rb = blk * 16 + lane / 4;
        const int lpc = ((lane % 4) - (rb >> 2)) & 3;
        const int c16 = rb % 16, l = rb / 16 % NL, tile = rb / 16 / NL;
        b_g[q] = (tile * SPT + c16 / B) * (kMaxN + 1);
        b_hi[q] = lpc >= 2;
        b_full[q] = l * limb_block_b + ((uint32_t)(c16 % B) << 6) + lpc * 16;
        b_tail[q] = l * limb_block_b + ((uint32_t)(c16 % B) << 5) + (lpc & 1) * 16;
    });
    // plane bases of the table whose full chunks are being staged (read from
    // LDS once per table, not once per chunk)
    const int8_t* b_base[PER];
    int b_table = -1;
    auto issue_b = [&](int8_t* dst, int i, int i2, uint32_t k0) {
        const bool tail = k0 == k_full;
        if (!tail && i != b_table) {
            static_for<PER>([&](auto q) { b_base[q] = s_bptr[b_g[q] + i]; });
            b_table = i;
        }
        static_for<PER>([&](auto q) {
            const int blk = wave + (int)decltype(q)::value * NW;  // wave-uniform
            if (B_BLOCKS % NW != 0 && blk >= B_BLOCKS) return;
            const int8_t* src = tail ? s_bptr[b_g[q] + (b_hi[q] ? i2 : i)] + (k0 * B + b_tail[q])
                                     : b_base[q] + (k0 * B + b_full[q]);
            load_to_lds(src, dst + blk * 1024);
        });
    };
    // this lane's 16 bytes of every table fragment of chunk (i, i2, k0): one
    // 32-bit lane offset per chunk, the fragments at uniform steps from it
    // (row tiles 16 chunk rows apart, limbs one limb block apart).  Every load
    // is in bounds, so no lane needs a branch: the tail partner i2 may be n,
    // and table index n is the all-zero table after the digits.  A row tile
    // past the slice (its MFMAs are skipped) loads nothing, a wave-uniform
    // branch: in c4's transpose sweep (400 rows = 7 + 6 + 6 + 6 tiles of 8
    // slots) that is 22 % of the fragment loads, in the forward 12.5 %.
    const uint32_t row_a = row_on[0] ? wg_row0 + wr * RT * 16 + cc : (uint32_t)cc;
    auto load_a = [&](v4i (&fa)[RT][NL], int i, int i2, uint32_t k0) {
        const bool tail = k0 == k_full;
        const uint32_t wshift = tail ? 5u : 6u;
        const int8_t* base = a.tabl + k0 * a.r_pad;
        const int t = (tail && kq >= 32) ? i2 : i;
        const uint32_t off = (uint32_t)t * table_stride + (row_a << wshift) + (uint32_t)(tail ? kq & 31 : kq);
        static_for<RT>([&](auto rt) {
            // wave-uniform: a row tile past the slice loads nothing (its
            // MFMAs are skipped, so its fragments may hold anything)
            if (!row_on[rt]) return;
            const uint32_t roff = (uint32_t)decltype(rt)::value * (16u << wshift);
            static_for<NL>([&](auto j) {
                fa[rt][j] = *reinterpret_cast<const v4i*>(base + (off + roff + (uint32_t)decltype(j)::value * limb_block_a));
            });
        });
    };

    auto first = [&](int& i, int& i2, uint32_t& k0) {
        i = next_i(0);
        if (k_full > 0) {
            i2 = n;
            k0 = 0;
        } else {
            i2 = i < n ? next_i(i + 1) : n;
            k0 = k_full;
        }
    };
    auto advance = [&](int& i, int& i2, uint32_t& k0) {
        if (k0 != k_full) {
            k0 += 64;
            if (k0 < k_full) return;
            i = next_i(i + 1);
            if (i < n) {
                k0 = 0;
                return;
            }
            if (!has_tail) return;
            i = next_i(0);
            i2 = i < n ? next_i(i + 1) : n;
            k0 = k_full;
            return;
        }
        i = i2 < n ? next_i(i2 + 1) : n;
        i2 = i < n ? next_i(i + 1) : n;
    };

    auto stage = [&](auto st) -> int8_t* {
        constexpr int X = decltype(st)::value % STG;
        if constexpr (X == 0) return s_st0;
        else if constexpr (X == 1) return s_st1;
        else return s_st2;
    };
    v4i fa0[RT][NL], fa1[RT][NL];
    // chunk c = (i, i2, k0): digit planes in stage ST mod STG, table fragments
    // in fa<ST mod 2>; c+1 = (j1, j12, q1); false after the last chunk
    int i, i2, j1, j12;
    uint32_t k0, q1;
    auto step = [&](auto ST) {
        constexpr int S = decltype(ST)::value;
        v4i(&fcur)[RT][NL] = S % 2 == 0 ? fa0 : fa1;
        v4i(&fnext)[RT][NL] = S % 2 == 0 ? fa1 : fa0;
        // this wave's loads of chunk c (staging and fragments) are done (with
        // three stages the staging of c+1 may stay in flight); after the
        // barrier every wave's staging of c is, and the stage of c-1 is free
        if (STG == 3 && j1 < n) wait_vmcnt<PER>(); else wait_vmcnt<0>();
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        int j2 = j1, j22 = j12;
        uint32_t q2 = q1;
        if (j1 < n) {
            load_a(fnext, j1, j12, q1);
            advance(j2, j22, q2);
            if constexpr (STG == 3) {
                if (j2 < n) issue_b(stage(std::integral_constant<int, S + 2>{}), j2, j22, q2);
            } else {
                issue_b(stage(std::integral_constant<int, S + 1>{}), j1, j12, q1);
            }
        }
        const int8_t* sb = stage(ST);
        v4i fb[CT][NL];
        static_for<CT>([&](auto ct) {
            static_for<NL>([&](auto l) {
                fb[ct][l] = *reinterpret_cast<const v4i*>(sb + ((wc * CT + ct) * NL + l) * 16 * kLdsStride +
                                                          lds_off(cc, kq));
            });
        });
        static_for<CT>([&](auto ct) {
            if (!chunk_on(on_mask[ct], i, i2, k0)) return;
            static_for<RT>([&](auto rt) {
                if (!row_on[rt]) return;
                static_for<NL>([&](auto j) {
                    static_for<NL>([&](auto l) {
                        acc[rt][ct][j + l] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fcur[rt][j], fb[ct][l],
                                                                                  acc[rt][ct][j + l], 0, 0, 0);
                    });
                });
            });
        });
        const bool more = j1 < n;
        i = j1;
        i2 = j12;
        k0 = q1;
        j1 = j2;
        j12 = j22;
        q1 = q2;
        return more;
    };
    first(i, i2, k0);
    if (i < n) {
        j1 = i;
        j12 = i2;
        q1 = k0;
        advance(j1, j12, q1);
        load_a(fa0, i, i2, k0);
        issue_b(s_st0, i, i2, k0);
        if constexpr (STG == 3) {
            if (j1 < n) issue_b(s_st1, j1, j12, q1);
            while (step(std::integral_constant<int, 0>{}) && step(std::integral_constant<int, 1>{}) &&
                   step(std::integral_constant<int, 2>{}) && step(std::integral_constant<int, 3>{}) &&
                   step(std::integral_constant<int, 4>{}) && step(std::integral_constant<int, 5>{})) {
            }
        } else {
            while (step(std::integral_constant<int, 0>{}) && step(std::integral_constant<int, 1>{})) {
            }
        }
    }
    uint32_t sc[RT][CT][4];
    mfma_scales<B, RT, CT, WR, WC>(a, cg, wg_row0, wg_tiles, sc);
    mfma_store<B, NL, RT, CT, WR, WC>(a, cg, wg_row0, wg_tiles, acc, sc);
}

template <int B, int NL, int RT, int CT, int WR, int WC, int STG, bool FWD>
__global__ __launch_bounds__(64 * WR * WC, 2) void k_dense_mfma_areg(MfmaSweepArgs a) {
    dense_mfma_areg_body<B, NL, RT, CT, WR, WC, STG, FWD>(a, blockIdx.x);
}

// many sectors' sweeps in one launch (same B, NL, tile and direction): the
// arguments of job j at jobs[j], its tiles [first[j], first[j+1]) of
// `total`.  Workgroup w walks tiles w, w + grid, w + 2 grid, ...: with one
// workgroup per tile that is one tile; as a persistent launch (a grid of a
// few workgroups per CU) it is every grid-th tile.  Every first[j] and the
// grid are multiples of 8, so a tile runs on the XCD its job-local index maps
// it to.  A barrier opens each tile: the waves of the previous one may still
// read the LDS stages the next one fills.
template <int B, int NL, int RT, int CT, int WR, int WC, int STG, bool FWD>
__global__ __launch_bounds__(64 * WR * WC, 2) void k_dense_mfma_areg_group(
    const MfmaSweepArgs* __restrict__ jobs, const uint32_t* __restrict__ first, uint32_t njobs, uint32_t total) {
    for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
        __syncthreads();
        const uint32_t j = group_job_at(first, njobs, t);
        const MfmaSweepArgs a = jobs[j];
        dense_mfma_areg_body<B, NL, RT, CT, WR, WC, STG, FWD>(a, t - first[j]);
    }
}

// Arguments and grid of one sweep launch of tile configuration (RT, CT, WR, WC).
template <int B, int RT, int CT, int WR, int WC>
static uint64_t mfma_args_cfg(const DenseSweep& s, const MfmaTable& t, const int8_t* lp, uint32_t* out,
                              const uint32_t* scale, const ModP& m, MfmaSweepArgs& a) {
    constexpr int SLOTS = WC * CT * (16 / B);
    constexpr int ROWS = WR * RT * 16;
    require(s.n <= (uint32_t)kMaxN, "MFMA sweep supports n <= 32");
    require(s.k < 16u, "MFMA sweep supports k < 16");
    a = MfmaSweepArgs{};
    a.inc = s.inc;
    require(a.inc != nullptr, "MFMA sweep needs the operator's incidence table");
    a.tabl = t.d;
    a.lp = lp;
    a.zero = t.zero;
    a.n = s.n;
    a.deg_out = s.fwd ? s.k - 1 : s.k;
    a.u_out = s.u_out;
    a.v_in = s.v_in;
    a.r_out = s.r_out;
    a.r_pad = t.rows_pad;
    a.k_in = s.k_in;
    a.k_pad = t.k_pad;
    require(limb_stride(s.k_in) == t.k_pad, "digit planes and table disagree on K padding");
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
    a.xcd_items = (a.col_groups * a.row_slices + 7) / 8;
    const uint64_t blocks = 8 * a.xcd_items;
    require(blocks < (1ULL << 31), "MFMA sweep grid too large");
    return blocks;
}

// kernel variants of a tile configuration: DMA 0 register-staged, 2 or 3
// direct-to-LDS stages, 4 + s table fragments in registers and s digit-plane
// stages
template <int B, int NL, int RT, int CT, int WR, int WC, int DMA = 0>
static void dense_mfma_cfg(const DenseSweep& s, const MfmaTable& t, const int8_t* lp, uint32_t* out,
                           const uint32_t* scale, const ModP& m, hipStream_t st) {
    MfmaSweepArgs a;
    const uint64_t blocks = mfma_args_cfg<B, RT, CT, WR, WC>(s, t, lp, out, scale, m, a);
#define PGW_MF(K, ...) hipLaunchKernelGGL((K<B, NL, RT, CT, WR, WC, __VA_ARGS__>), dim3((unsigned)blocks), \
                                          dim3(64 * WR * WC), 0, st, a)
    if constexpr (DMA >= 4) {
        if (s.fwd) PGW_MF(k_dense_mfma_areg, DMA - 4, true); else PGW_MF(k_dense_mfma_areg, DMA - 4, false);
    } else if constexpr (DMA) {
        if (s.fwd) PGW_MF(k_dense_mfma_dma, DMA, true); else PGW_MF(k_dense_mfma_dma, DMA, false);
    } else {
        if (s.fwd) PGW_MF(k_dense_mfma, true); else PGW_MF(k_dense_mfma, false);
    }
#undef PGW_MF
    PGW_CHECK_LAUNCH("MFMA sweep");
}

// Grouped sweeps use the register-A configurations (7, 8, 9).  prepare: the
// job's arguments and its block count rounded up to a multiple of 8.
template <int B, int NL>
static uint64_t mfma_group_prepare_t(int cfg, const DenseSweep& s, const MfmaTable& t, const int8_t* lp,
                                     uint32_t* out, const uint32_t* scale, const ModP& m, MfmaSweepArgs& a) {
    constexpr int CT = NL == 4 ? 2 : 4;
    uint64_t blocks = 0;
    switch (cfg) {
        case 7: blocks = mfma_args_cfg<B, 2, CT, 8, 1>(s, t, lp, out, scale, m, a); break;
        case 8: case 9: blocks = mfma_args_cfg<B, 2, CT, 4, 1>(s, t, lp, out, scale, m, a); break;
        default: throw Error(PGW_ERR_UNSUPPORTED, "grouped MFMA sweeps use tile configurations 7-9");
    }
    return (blocks + 7) / 8 * 8;
}

template <int B, int NL>
static void mfma_group_launch_t(int cfg, bool fwd, const MfmaSweepArgs* jobs, const uint32_t* first,
                                uint32_t njobs, uint64_t blocks, bool persistent, hipStream_t st) {
    constexpr int CT = NL == 4 ? 2 : 4;
    require(blocks < (1ULL << 31) && blocks % 8 == 0, "grouped MFMA sweep grid too large or unaligned");
    // one workgroup per tile, or persistent: two workgroups per CU
    // (configuration 7's eight waves: one) walking all the group's tiles
    const unsigned grid = persistent ? (unsigned)std::min<uint64_t>(blocks, persistent_grid(cfg == 7 ? 1 : 2))
                                     : (unsigned)blocks;
#define PGW_MG(WR, STG, F)                                                                          \
    hipLaunchKernelGGL((k_dense_mfma_areg_group<B, NL, 2, CT, WR, 1, STG, F>), dim3(grid),           \
                       dim3(64 * WR), 0, st, jobs, first, njobs, (uint32_t)blocks)
    switch (cfg) {
        case 7: if (fwd) PGW_MG(8, 2, true); else PGW_MG(8, 2, false); break;
        case 8: if (fwd) PGW_MG(4, 2, true); else PGW_MG(4, 2, false); break;
        case 9: if (fwd) PGW_MG(4, 3, true); else PGW_MG(4, 3, false); break;
        default: throw Error(PGW_ERR_UNSUPPORTED, "grouped MFMA sweeps use tile configurations 7-9");
    }
#undef PGW_MG
    PGW_CHECK_LAUNCH("grouped MFMA sweep");
}

template <int B, int NL>
static void dense_mfma_pick(int cfg, const DenseSweep& s, const MfmaTable& t, const int8_t* lp,
                            uint32_t* out, const uint32_t* scale, const ModP& m, hipStream_t st) {
    if constexpr (NL == 4) {
        switch (cfg) {
            case 0: dense_mfma_cfg<B, NL, 2, 2, 2, 2>(s, t, lp, out, scale, m, st); return;
            case 1: dense_mfma_cfg<B, NL, 2, 2, 4, 2>(s, t, lp, out, scale, m, st); return;
            case 2: dense_mfma_cfg<B, NL, 2, 2, 2, 4>(s, t, lp, out, scale, m, st); return;
            case 3: dense_mfma_cfg<B, NL, 2, 2, 4, 2, 3>(s, t, lp, out, scale, m, st); return;
            case 4: dense_mfma_cfg<B, NL, 2, 2, 8, 1, 2>(s, t, lp, out, scale, m, st); return;
            case 5: dense_mfma_cfg<B, NL, 2, 2, 2, 4, 3>(s, t, lp, out, scale, m, st); return;
            case 6: dense_mfma_cfg<B, NL, 2, 2, 2, 2, 2>(s, t, lp, out, scale, m, st); return;
            case 7: dense_mfma_cfg<B, NL, 2, 2, 8, 1, 6>(s, t, lp, out, scale, m, st); return;
            case 8: dense_mfma_cfg<B, NL, 2, 2, 4, 1, 6>(s, t, lp, out, scale, m, st); return;
            default: dense_mfma_cfg<B, NL, 2, 2, 4, 1, 7>(s, t, lp, out, scale, m, st); return;
        }
    } else {
        switch (cfg) {
            case 0: dense_mfma_cfg<B, NL, 2, 4, 2, 2>(s, t, lp, out, scale, m, st); return;
            case 1: dense_mfma_cfg<B, NL, 2, 4, 4, 2>(s, t, lp, out, scale, m, st); return;
            case 2: dense_mfma_cfg<B, NL, 4, 2, 2, 4>(s, t, lp, out, scale, m, st); return;
            case 3: dense_mfma_cfg<B, NL, 2, 4, 4, 2, 3>(s, t, lp, out, scale, m, st); return;
            case 4: dense_mfma_cfg<B, NL, 2, 4, 8, 1, 2>(s, t, lp, out, scale, m, st); return;
            case 5: dense_mfma_cfg<B, NL, 4, 2, 2, 4, 3>(s, t, lp, out, scale, m, st); return;
            case 6: dense_mfma_cfg<B, NL, 2, 4, 2, 2, 2>(s, t, lp, out, scale, m, st); return;
            case 7: dense_mfma_cfg<B, NL, 2, 4, 8, 1, 6>(s, t, lp, out, scale, m, st); return;
            case 8: dense_mfma_cfg<B, NL, 2, 4, 4, 1, 6>(s, t, lp, out, scale, m, st); return;
            default: dense_mfma_cfg<B, NL, 2, 4, 4, 1, 7>(s, t, lp, out, scale, m, st); return;
        }
    }
}

}  // namespace pgw

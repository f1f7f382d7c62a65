// pgw_gram_mfma.hip -- the block projection X^T Y (and Y^T Y or X^T X) on
// int8 MFMA: the residues of both blocks are split into balanced base-256
// digits in registers (the representation of pgw_fold.h), the digit products
// of each shift accumulate in int32 over chunks of 64 rows, and the shift
// sums are folded back to residues before any accumulator can pass 2^31.
// It writes the same per-workgroup partials as the vector-ALU projection,
// which k_gram_finish sums and reduces.  Replaces the role of
// dense_gemm_TN_chunked3D_offset_u32 (cuda/prym_cuda_helpers.h:107-137),
// whose u32 sums need terms * p^2 < 2^32.
//
// Operand layout of v_mfma_i32_16x16x64_i8: lane L holds A[L % 16][16 (L / 16)
// .. +15] and B[16 (L / 16) .. +15][L % 16]; for a Gram product both are
// "column L % 16 of 16 consecutive rows", so one set of digit registers per
// block serves as A and as B.  For b < 16 the 16 MFMA rows/columns are
// 16 / b row groups of b columns: row group g reads rows 64 g .. 64 g + 63 of
// the chunk, and only the diagonal blocks (same group) of the 16 x 16 result
// are kept, each the Gram product of its own 64 rows.
#include <hip/hip_runtime.h>

#include "pgw_fold.h"
#include "pgw_kernels.h"

namespace pgw {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kGramMfmaWaves = 4;
// chunks between folds: a chunk adds to each shift at most NL MFMAs of 64
// products |d e| <= 2^14, so 256 chunks stay within 256 * 4 * 2^20 = 2^30
constexpr int kGramMfmaFlush = 256;

// digit j of the 16 values v[t], as int8 bytes t = 0..15 of out[j]
template <int NL>
__device__ __forceinline__ void split_digits16(const uint32_t (&v)[16], uint32_t p, const DigitSpec& ds,
                                               v4i (&out)[NL]) {
    uint32_t u[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) u[t] = (uint32_t)(digits_biased(v[t], p, ds) ^ ds.bias);
#pragma unroll
    for (int j = 0; j < NL; ++j) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            out[j][q] = (int)(((u[4 * q] >> (8 * j)) & 0xffu) | (((u[4 * q + 1] >> (8 * j)) & 0xffu) << 8) |
                              (((u[4 * q + 2] >> (8 * j)) & 0xffu) << 16) |
                              (((u[4 * q + 3] >> (8 * j)) & 0xffu) << 24));
        }
    }
}

template <int NL>
__device__ __forceinline__ void fold_into(v4i (&acc)[2 * NL - 1], uint32_t (&r)[4], const ModP& m,
                                          const ShiftWeights& sw) {
    constexpr int NS = 2 * NL - 1;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        int32_t x[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) x[s] = acc[s][v];
        r[v] = addmod(r[v], fold_digit_shifts<NL>(x, m, sw.w), m.p);
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[s] = v4i{0, 0, 0, 0};
}

template <int B, int NL>
__global__ __launch_bounds__(64 * kGramMfmaWaves) void k_gram_mfma(const uint32_t* __restrict__ x,
                                                                   const uint32_t* __restrict__ y,
                                                                   uint64_t rows, int mode,
                                                                   uint64_t* __restrict__ part, ModP m,
                                                                   ShiftWeights sw) {
    constexpr int G = 16 / B;   // row groups per MFMA
    constexpr int RM = 64 * G;  // rows per MFMA chunk
    constexpr int NS = 2 * NL - 1;
    constexpr int P = B * (B + 1) / 2;
    const DigitSpec ds = digit_spec(NL);
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int i16 = lane % 16, grp = i16 / B, col = i16 % B, kb = (lane / 16) * 16;
    const uint32_t nb = gridDim.x;

    v4i acc0[NS], acc1[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) acc0[s] = acc1[s] = v4i{0, 0, 0, 0};
    uint32_t r0[4] = {0, 0, 0, 0}, r1[4] = {0, 0, 0, 0};
    int since = 0;
    const uint64_t stride = (uint64_t)nb * kGramMfmaWaves * RM;
    for (uint64_t base = ((uint64_t)blockIdx.x * kGramMfmaWaves + wave) * RM; base < rows; base += stride) {
        uint32_t xv[16], yv[16];
        const uint64_t r_lane = base + (uint64_t)grp * 64 + kb;
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const uint64_t row = r_lane + t;
            const bool ok = row < rows;
            xv[t] = ok ? x[row * B + col] : 0u;
            yv[t] = ok ? y[row * B + col] : 0u;
        }
        v4i dx[NL], dy[NL];
        split_digits16<NL>(xv, m.p, ds, dx);
        split_digits16<NL>(yv, m.p, ds, dy);
#pragma unroll
        for (int j = 0; j < NL; ++j)
#pragma unroll
            for (int l = 0; l < NL; ++l)
                acc0[j + l] = __builtin_amdgcn_mfma_i32_16x16x64_i8(dx[j], dy[l], acc0[j + l], 0, 0, 0);
        if (mode == 1) {
#pragma unroll
            for (int j = 0; j < NL; ++j)
#pragma unroll
                for (int l = 0; l < NL; ++l)
                    acc1[j + l] = __builtin_amdgcn_mfma_i32_16x16x64_i8(dy[j], dy[l], acc1[j + l], 0, 0, 0);
        } else if (mode == 2) {
#pragma unroll
            for (int j = 0; j < NL; ++j)
#pragma unroll
                for (int l = 0; l < NL; ++l)
                    acc1[j + l] = __builtin_amdgcn_mfma_i32_16x16x64_i8(dx[j], dx[l], acc1[j + l], 0, 0, 0);
        }
        if (++since == kGramMfmaFlush) {
            fold_into<NL>(acc0, r0, m, sw);
            if (mode) fold_into<NL>(acc1, r1, m, sw);
            since = 0;
        }
    }
    fold_into<NL>(acc0, r0, m, sw);
    if (mode) fold_into<NL>(acc1, r1, m, sw);

    // the diagonal blocks of the lane's four results (row 4 (L / 16) + v,
    // column L % 16) summed per entry (i, j) of the b x b product: at most
    // 4 waves x 4 groups residues, < 2^35
    __shared__ unsigned long long s_sum[2][B * B];
    for (int e = threadIdx.x; e < 2 * B * B; e += 64 * kGramMfmaWaves) (&s_sum[0][0])[e] = 0;
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int i = 4 * (lane / 16) + v, j = lane % 16;
        if (i / B != j / B) continue;
        const int e = (i % B) * B + (j % B);
        atomicAdd(&s_sum[0][e], (unsigned long long)r0[v]);
        if (mode) atomicAdd(&s_sum[1][e], (unsigned long long)r1[v]);
    }
    __syncthreads();
    // packed upper triangle, as k_gram_partial writes it: part[((mat P +
    // pair) nblocks + block) * 2 + {0, 1}] = (low, high) 32-bit words of the sum
    for (int e = threadIdx.x; e < P * (mode ? 2 : 1); e += 64 * kGramMfmaWaves) {
        const int mat = e / P, pr = e % P;
        int i = 0, rem = pr;
        while (rem >= B - i) {
            rem -= B - i;
            ++i;
        }
        const unsigned long long s = s_sum[mat][i * B + i + rem];
        uint64_t* o = part + ((uint64_t)e * nb + blockIdx.x) * 2;
        o[0] = s & 0xffffffffULL;
        o[1] = s >> 32;
    }
}

}  // namespace

unsigned gram_mfma_grid(uint64_t rows, uint32_t b, unsigned max_grid) {
    const uint64_t per = (uint64_t)kGramMfmaWaves * 64 * (16 / b);
    return (unsigned)std::min<uint64_t>(max_grid, std::max<uint64_t>(1, (rows + per - 1) / per));
}

void launch_gram_mfma_partial(const uint32_t* x, const uint32_t* y, uint64_t rows, uint32_t b, int mode,
                              uint64_t* part, unsigned grid, const ModP& m, hipStream_t st) {
    require(b == 4 || b == 8 || b == 16, "MFMA Gram projection needs b in {4, 8, 16}");
    ShiftWeights sw{};
    uint64_t w = 1;
    for (int q = 0; q < 8; ++q) {
        sw.w[q] = (uint32_t)w;
        w = (w * 256) % m.p;
    }
    const bool two = m.p <= 65536;  // balanced two-digit representation
#define PGW_GM(BB, NL)                                                                                 \
    hipLaunchKernelGGL((k_gram_mfma<BB, NL>), dim3(grid), dim3(64 * kGramMfmaWaves), 0, st, x, y, rows, \
                       mode, part, m, sw)
    if (b == 16) { if (two) PGW_GM(16, 2); else PGW_GM(16, 4); }
    else if (b == 8) { if (two) PGW_GM(8, 2); else PGW_GM(8, 4); }
    else { if (two) PGW_GM(4, 2); else PGW_GM(4, 4); }
#undef PGW_GM
    PGW_CHECK_LAUNCH("MFMA gram partial");
}

}  // namespace pgw

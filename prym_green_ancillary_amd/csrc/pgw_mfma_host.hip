// pgw_mfma_host.hip -- host side and digit-plane kernel of the int8-limb MFMA
// table contraction: table digits, limb split, launch dispatch by block width.
#include "pgw_mfma_sweep.h"

#include <cstdlib>

namespace pgw {

static inline uint32_t round_up(uint32_t x, uint32_t a) { return (x + a - 1) / a * a; }

// byte of (row, k) in a chunk-major block of `rows` rows and K_pad bytes
__host__ __device__ inline uint64_t chunk_major_offset(uint64_t row, uint32_t k, uint64_t rows,
                                                       uint32_t k_pad) {
    const uint32_t k_full = k_pad / 64 * 64;
    return k < k_full ? (uint64_t)(k / 64) * rows * 64 + row * 64 + (k % 64)
                      : (uint64_t)k_full * rows + row * 32 + (k - k_full);
}

uint32_t mfma_limbs(uint32_t p) { return p <= 65536u ? 2u : 4u; }
bool mfma_block_supported(uint32_t b) { return b == 4 || b == 8 || b == 16; }
// K is padded to a multiple of 32 and stored chunk-major: for each 64-byte K
// chunk the rows (table rows, or panel columns) follow each other, 64 bytes
// each, and a 32-byte tail chunk (K_pad % 64 == 32) follows with 32 bytes per
// row.  A 16-row slice of a chunk is one aligned, contiguous 1 KB block.
uint32_t limb_stride(uint32_t k_in) { return round_up(k_in, 32); }

uint64_t limb_plane_bytes(uint64_t v, uint32_t k_in, uint32_t b, uint32_t nl) {
    return 2ULL * v * nl * b * limb_stride(k_in) + 256;
}

void build_mfma_table(MfmaTable& t, uint32_t n, uint32_t rows, uint32_t k_in, uint32_t p,
                      const uint32_t* src, bool transpose_src) {
    // !transpose_src: src[i][row][k] (rows x k_in); transpose_src: src[i][k][row]
    const uint32_t nl = mfma_limbs(p);
    t.rows = rows;
    t.k_in = k_in;
    t.rows_pad = round_up(rows, 16);
    t.k_pad = round_up(k_in, 32);
    const DigitSpec ds = digit_spec(nl);
    std::vector<int8_t> host((size_t)n * nl * t.rows_pad * t.k_pad, 0);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t r = 0; r < rows; ++r)
            for (uint32_t k = 0; k < k_in; ++k) {
                uint32_t v = transpose_src ? src[((size_t)i * k_in + k) * rows + r]
                                           : src[((size_t)i * rows + r) * k_in + k];
                uint64_t u = digits_biased(v, p, ds);
                for (uint32_t j = 0; j < nl; ++j)
                    host[((size_t)i * nl + j) * t.rows_pad * t.k_pad +
                         chunk_major_offset(r, k, t.rows_pad, t.k_pad)] =
                        (int8_t)(uint8_t)(((u >> (8 * j)) & 0xff) ^ 0x80);
            }
    const size_t digits = host.size();
    host.resize(digits + mfma_zero_bytes(nl, t.rows_pad, t.k_pad), 0);
    PGW_HIP(hipMalloc(&t.d, host.size()));
    PGW_HIP(hipMemcpy(t.d, host.data(), host.size(), hipMemcpyHostToDevice));
    t.zero = t.d + digits;
}

// Incidences of every output panel, computed once per operator so that a
// sweep workgroup reads its panels' codes in one coalesced pass instead of
// unranking and ranking subsets through chains of dependent lookups.
// (cuda/cuprym_cyclic.cu:631-727 builds the same incidences and signs.)
std::vector<uint32_t> mfma_incidence_table(uint32_t n, uint32_t deg_out, bool fwd, uint64_t u_out,
                                           uint64_t v_in) {
    require(n <= 32 && (fwd || deg_out >= 1), "incidence table: bad subset degree");
    require(v_in < (1ULL << 31), "incidence table: neighbour ranks need 31 bits");
    struct Pascal {
        uint64_t c[33][33] = {};
        uint64_t operator()(int a, int b) const { return (a < 0 || b < 0 || b > a) ? 0ULL : c[a][b]; }
    } C;
    for (int a = 0; a <= 32; ++a) {
        C.c[a][0] = 1;
        for (int b = 1; b <= a; ++b) C.c[a][b] = C.c[a - 1][b - 1] + (b < a ? C.c[a - 1][b] : 0);
    }
    const uint64_t w = (uint64_t)n + 1;
    std::vector<uint32_t> t(u_out * w);
    for (uint64_t u = 0; u < u_out; ++u) {
        const uint64_t mask = unrank_mask(u, (int)n, (int)deg_out, C);
        uint32_t* row = t.data() + u * w;
        uint32_t on = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const uint64_t bit = 1ULL << i;
            uint32_t code = 0xffffffffu;
            if (fwd ? !(mask & bit) : (mask & bit)) {
                const uint32_t neg = (uint32_t)__builtin_popcountll(mask & (bit - 1)) & 1u;
                const uint64_t nb = fwd ? rank_mask(mask | bit, (int)n, (int)deg_out + 1, C)
                                        : rank_mask(mask & ~bit, (int)n, (int)deg_out - 1, C);
                require(nb < v_in, "incidence table: neighbour rank out of range");
                code = (uint32_t)nb | (neg << 31);
                on |= 1u << i;
            }
            row[1 + i] = code;
        }
        row[0] = on;
    }
    return t;
}

// ---------------------------------------------------------------------------
// digit planes of a block
// ---------------------------------------------------------------------------

// one thread per (panel v, 16-row group kg, column c): 16 words of a column in,
// 16 digit bytes per limb and sign out (one 16-byte store each)
template <int B, int NL>
__device__ __forceinline__ void limb_split_one(const uint32_t* __restrict__ in,
                                               const uint32_t* __restrict__ scale,
                                               int8_t* __restrict__ lp, uint64_t v_count,
                                               uint32_t k_in, uint32_t k_stride, const ModP& m,
                                               uint64_t tid) {
    const uint32_t groups = k_stride / 16;
    const uint64_t total = v_count * groups * B;
    if (tid >= total) return;
    const int c = (int)(tid % B);
    const uint64_t rest = tid / B;
    const uint32_t kg = (uint32_t)(rest % groups);
    const uint64_t v = rest / groups;
    const DigitSpec ds = digit_spec(NL);
    uint32_t w[2][NL][4];
    static_for<2>([&](auto s) {
        static_for<NL>([&](auto j) { static_for<4>([&](auto q) { w[s][j][q] = 0; }); });
    });
    static_for<16>([&](auto tt) {
        const uint32_t k = kg * 16 + tt;
        uint32_t x = 0;
        if (k < k_in) {
            const uint64_t row = v * k_in + k;
            x = in[row * B + c];
            if (scale) {
                // x * scale < 2^62
                x = m.p == kM31 ? red63_m31((uint64_t)x * scale[row]) : mulmod(x, scale[row], m);
            }
        }
        const uint32_t xn = x ? m.p - x : 0u;
        const uint64_t u0 = digits_biased(x, m.p, ds), u1 = digits_biased(xn, m.p, ds);
        static_for<NL>([&](auto j) {
            w[0][j][tt / 4] |= (uint32_t)((((u0 >> (8 * j)) & 0xff) ^ 0x80) << (8 * (tt % 4)));
            w[1][j][tt / 4] |= (uint32_t)((((u1 >> (8 * j)) & 0xff) ^ 0x80) << (8 * (tt % 4)));
        });
    });
    const uint64_t within = chunk_major_offset((uint64_t)c, kg * 16, B, k_stride);
    static_for<2>([&](auto s) {
        static_for<NL>([&](auto j) {
            int8_t* dst = lp + (((uint64_t)s * v_count + v) * NL + j) * B * k_stride + within;
            *reinterpret_cast<uint4*>(dst) = make_uint4(w[s][j][0], w[s][j][1], w[s][j][2], w[s][j][3]);
        });
    });
}

template <int B, int NL>
__global__ __launch_bounds__(256) void k_limb_split(const uint32_t* __restrict__ in,
                                                    const uint32_t* __restrict__ scale,
                                                    int8_t* __restrict__ lp, uint64_t v_count,
                                                    uint32_t k_in, uint32_t k_stride, ModP m) {
    limb_split_one<B, NL>(in, scale, lp, v_count, k_in, k_stride, m,
                          (uint64_t)blockIdx.x * 256 + threadIdx.x);
}

// many jobs' digit splits in one launch: job j's blocks [first[j], first[j+1])
template <int B, int NL>
__global__ __launch_bounds__(256) void k_limb_split_group(const LimbJob* __restrict__ jobs,
                                                          const uint32_t* __restrict__ first,
                                                          uint32_t njobs) {
    const uint32_t j = group_job(first, njobs);
    const LimbJob& J = jobs[j];
    limb_split_one<B, NL>(J.in, J.scale, J.lp, J.v_count, J.k_in, J.k_stride, J.m,
                          (uint64_t)(blockIdx.x - first[j]) * 256 + threadIdx.x);
}

static inline unsigned grid_for(uint64_t threads, unsigned block = 256) {
    uint64_t g = (threads + block - 1) / block;
    return (unsigned)std::max<uint64_t>(1, g);
}

void launch_limb_split(const uint32_t* in, const uint32_t* scale, int8_t* lp, uint64_t v,
                       uint32_t k_in, uint32_t b, uint32_t nl, const ModP& m, hipStream_t st) {
    const uint32_t ks = limb_stride(k_in);
    const uint64_t total = v * (ks / 16) * b;
    if (total == 0) return;
    dim3 grid(grid_for(total)), blk(256);
#define PGW_LS(BB, NN) hipLaunchKernelGGL((k_limb_split<BB, NN>), grid, blk, 0, st, in, scale, lp, v, k_in, ks, m)
    if (nl == 2) {
        if (b == 4) PGW_LS(4, 2); else if (b == 8) PGW_LS(8, 2); else if (b == 16) PGW_LS(16, 2);
        else throw Error(PGW_ERR_UNSUPPORTED, "MFMA path supports b in {4, 8, 16}");
    } else {
        if (b == 4) PGW_LS(4, 4); else if (b == 8) PGW_LS(8, 4); else if (b == 16) PGW_LS(16, 4);
        else throw Error(PGW_ERR_UNSUPPORTED, "MFMA path supports b in {4, 8, 16}");
    }
#undef PGW_LS
    PGW_CHECK_LAUNCH("limb split");
}

uint64_t limb_group_blocks(const LimbJob& j, uint32_t b) {
    return (j.v_count * (j.k_stride / 16) * b + 255) / 256;
}

void launch_limb_split_group(const LimbJob* jobs, const uint32_t* first, uint32_t njobs,
                             uint64_t blocks, uint32_t b, uint32_t nl, hipStream_t st) {
    if (blocks == 0) return;
    require(blocks < (1ULL << 31), "grouped digit split grid too large");
    const dim3 grid((unsigned)blocks), blk(256);
#define PGW_LG(BB, NN) hipLaunchKernelGGL((k_limb_split_group<BB, NN>), grid, blk, 0, st, jobs, first, njobs)
    if (nl == 2) {
        if (b == 4) PGW_LG(4, 2); else if (b == 8) PGW_LG(8, 2); else if (b == 16) PGW_LG(16, 2);
        else throw Error(PGW_ERR_UNSUPPORTED, "MFMA path supports b in {4, 8, 16}");
    } else {
        if (b == 4) PGW_LG(4, 4); else if (b == 8) PGW_LG(8, 4); else if (b == 16) PGW_LG(16, 4);
        else throw Error(PGW_ERR_UNSUPPORTED, "MFMA path supports b in {4, 8, 16}");
    }
#undef PGW_LG
    PGW_CHECK_LAUNCH("grouped limb split");
}

// ---------------------------------------------------------------------------
int mfma_config_count() { return 10; }

uint32_t persistent_grid(uint32_t per_cu) {
    int dev = 0, cus = 0;
    PGW_HIP(hipGetDevice(&dev));
    PGW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return std::max<uint32_t>(8, (uint32_t)cus * per_cu / 8 * 8);
}

void launch_dense_mfma_b4(int, uint32_t, const DenseSweep&, const MfmaTable&, const int8_t*,
                          uint32_t*, const uint32_t*, const ModP&, hipStream_t);
void launch_dense_mfma_b8(int, uint32_t, const DenseSweep&, const MfmaTable&, const int8_t*,
                          uint32_t*, const uint32_t*, const ModP&, hipStream_t);
void launch_dense_mfma_b16(int, uint32_t, const DenseSweep&, const MfmaTable&, const int8_t*,
                           uint32_t*, const uint32_t*, const ModP&, hipStream_t);

#define PGW_GROUP_DECL(BB)                                                                        \
    uint64_t mfma_group_prepare_b##BB(int, uint32_t, const DenseSweep&, const MfmaTable&,          \
                                      const int8_t*, uint32_t*, const uint32_t*, const ModP&,       \
                                      MfmaSweepArgs&);                                              \
    void mfma_group_launch_b##BB(int, uint32_t, bool, const MfmaSweepArgs*, const uint32_t*,        \
                                 uint32_t, uint64_t, bool, hipStream_t);
PGW_GROUP_DECL(4)
PGW_GROUP_DECL(8)
PGW_GROUP_DECL(16)
#undef PGW_GROUP_DECL

uint64_t mfma_group_prepare(int cfg, uint32_t b, uint32_t nl, const DenseSweep& s, const MfmaTable& t,
                            const int8_t* lp, uint32_t* out, const uint32_t* scale, const ModP& m,
                            MfmaSweepArgs& a) {
    if (b == 16) return mfma_group_prepare_b16(cfg, nl, s, t, lp, out, scale, m, a);
    if (b == 8) return mfma_group_prepare_b8(cfg, nl, s, t, lp, out, scale, m, a);
    if (b == 4) return mfma_group_prepare_b4(cfg, nl, s, t, lp, out, scale, m, a);
    throw Error(PGW_ERR_UNSUPPORTED, "MFMA path supports b in {4, 8, 16}");
}

void mfma_group_launch(int cfg, uint32_t b, uint32_t nl, bool fwd, const MfmaSweepArgs* jobs,
                       const uint32_t* first, uint32_t njobs, uint64_t blocks, bool persistent, hipStream_t st) {
    if (b == 16) mfma_group_launch_b16(cfg, nl, fwd, jobs, first, njobs, blocks, persistent, st);
    else if (b == 8) mfma_group_launch_b8(cfg, nl, fwd, jobs, first, njobs, blocks, persistent, st);
    else if (b == 4) mfma_group_launch_b4(cfg, nl, fwd, jobs, first, njobs, blocks, persistent, st);
    else throw Error(PGW_ERR_UNSUPPORTED, "MFMA path supports b in {4, 8, 16}");
}

void launch_dense_mfma(const DenseSweep& s, const MfmaTable& t, const int8_t* lp, uint32_t* out,
                       const uint32_t* scale, uint32_t b, uint32_t nl, const ModP& m,
                       hipStream_t st, int cfg) {
    if (cfg < 0 || cfg >= mfma_config_count()) cfg = 8;
    if (b == 16) launch_dense_mfma_b16(cfg, nl, s, t, lp, out, scale, m, st);
    else if (b == 8) launch_dense_mfma_b8(cfg, nl, s, t, lp, out, scale, m, st);
    else if (b == 4) launch_dense_mfma_b4(cfg, nl, s, t, lp, out, scale, m, st);
    else throw Error(PGW_ERR_UNSUPPORTED, "MFMA path supports b in {4, 8, 16}");
}

}  // namespace pgw

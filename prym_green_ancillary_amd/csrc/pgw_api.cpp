// pgw_api.cpp -- the C ABI of libpgw: contexts, operators, the block-apply
// chain, the Gram projection, the sequence loop and kernel-vector recovery.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "pgw_kernels.h"

using namespace pgw;

namespace {

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return PGW_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return PGW_ERR_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return PGW_ERR_INVALID;
    }
}

struct DeviceArena {
    std::vector<void*> ptrs;
    uint64_t bytes = 0;
    template <typename T>
    T* alloc(uint64_t count) {
        if (count == 0) count = 1;
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess)
            throw Error(PGW_ERR_MEMORY, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                            " bytes failed: " + hipGetErrorString(e));
        ptrs.push_back(p);
        bytes += count * sizeof(T);
        return (T*)p;
    }
    template <typename T>
    T* upload(const T* host, uint64_t count) {
        T* d = alloc<T>(count);
        if (count) PGW_HIP(hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice));
        return d;
    }
    template <typename T>
    T* upload(const std::vector<T>& v) { return upload(v.data(), (uint64_t)v.size()); }
    void release() {
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
        bytes = 0;
    }
};

PhiDevice upload_phi(DeviceArena& A, const PhiHost& h) {
    PhiDevice d;
    d.nI = h.nI; d.nB = h.nB; d.nA = h.nA; d.r = h.r;
    d.rows = h.row_rank.size();
    d.cols = h.col_rank.size();
    d.row_subsets = h.row_subsets;
    d.col_subsets = h.col_subsets;
    d.row_width = h.row_width;
    d.col_width = h.col_width;
    d.mu = A.upload(h.mu);
    d.wV = A.upload(h.wV);
    d.wA0 = A.upload(h.wA0);
    d.wA1 = A.upload(h.wA1);
    d.a0_off = A.upload(h.a0_off);
    d.a0_idx = A.upload(h.a0_idx);
    d.a1_off = A.upload(h.a1_off);
    d.a1_idx = A.upload(h.a1_idx);
    d.row_off = A.upload(h.row_off);
    d.row_rank = A.upload(h.row_rank);
    d.row_beta = A.upload(h.row_beta);
    d.col_off = A.upload(h.col_off);
    d.col_rank = A.upload(h.col_rank);
    d.col_a = A.upload(h.col_a);
    d.row_inc = A.upload(h.row_inc_code);
    d.row_inc_rank = A.upload(h.row_inc_rank);
    d.col_inc = A.upload(h.col_inc_code);
    d.col_inc_rank = A.upload(h.col_inc_rank);
    return d;
}

}  // namespace

// ---------------------------------------------------------------------------

struct pgw_op {
    pgw_ctx* ctx = nullptr;
    uint32_t kind = 0, p = 0, b = 0, method = PGW_METHOD_PLAIN;
    ModP m{};
    uint64_t rows = 0, cols = 0, mid = 0;
    DeviceArena arena;
    uint32_t* d_dr = nullptr;
    uint32_t* d_dc = nullptr;
    uint32_t* d_negdc = nullptr;
    pgw_rowmix_pair* d_pairs = nullptr;
    uint32_t rounds = 0;
    uint64_t pair_count = 0;
    int64_t unpaired = -1;
    uint32_t* d_scaled = nullptr;  // cols x b
    uint32_t* d_mid = nullptr;     // mid x b
    uint32_t* d_target = nullptr;  // rows x b
    uint64_t* d_gram = nullptr;
    // synthetic
    uint32_t n = 0, k = 0, M = 0, Mp = 0;
    uint64_t u_src = 0, u_tgt = 0;
    uint32_t* d_tab[4][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    uint32_t k4[4] = {0, 0, 0, 0};
    MfmaTable mt[4];
    int mfma_cfg[4] = {1, 1, 1, 1};  // tile configuration per sweep (pgw_operator_autotune)
    uint32_t nl = 0;
    int8_t* d_lp_src = nullptr;
    int8_t* d_lp_mid = nullptr;
    // cyclic eliminated
    PhiDevice dm, d2;
    uint32_t* d_zoff = nullptr;
    uint32_t* d_zrank = nullptr;
    uint16_t* d_zkappa = nullptr;
    uint32_t* d_R = nullptr;
    uint32_t* d_Bk = nullptr;
    uint32_t n_v = 0, kdim = 0;
    uint64_t x_cols = 0, z_count = 0;
    uint32_t* d_t = nullptr;
    uint32_t* d_y = nullptr;
    uint32_t* d_h = nullptr;
    uint32_t* d_tadj = nullptr;
    // timing of the sweep kernels (pgw_bench_apply)
    bool profile = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> sweep_events;
};

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
    return s;
}

void check_mfma_bound(const DenseSweep& s, uint32_t nl) {
    // int32 accumulators: incidences * K_in * (pairs per shift <= NL) * 2^14 < 2^31
    const uint64_t inc = s.fwd ? (s.n - s.k + 1) : s.k;
    require(inc * s.k_in * nl * (1ULL << 14) < (1ULL << 31),
            "int8-limb accumulation bound exceeded for this operator (incidences*K*limbs >= 2^17)",
            PGW_ERR_UNSUPPORTED);
}

struct SweepTimer {
    pgw_op* op;
    hipStream_t st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SweepTimer(pgw_op* o, hipStream_t s) : op(o), st(s) {
        if (op->profile) {
            PGW_HIP(hipEventCreate(&e0));
            PGW_HIP(hipEventCreate(&e1));
            PGW_HIP(hipEventRecord(e0, st));
        }
    }
    void done() {
        if (op->profile) {
            PGW_HIP(hipEventRecord(e1, st));
            op->sweep_events.emplace_back(e0, e1);
        }
    }
};

// in: u32 block (plain) ; the MFMA path consumes the digit planes prepared by
// the caller in op->d_lp_src / d_lp_mid
void run_sweep(pgw_op* op, int which, const uint32_t* in_u32, const int8_t* in_lp, uint32_t* out,
               const uint32_t* scale, hipStream_t st) {
    const DenseSweep s = sweep_desc(op, which);
    SweepTimer t(op, st);
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
            launch_limb_split(x, op->d_dc, op->d_lp_src, op->u_src, op->M, op->b, op->nl, m, st);
            run_sweep(op, 0, nullptr, op->d_lp_src, op->d_mid, nullptr, st);
            launch_limb_split(op->d_mid, nullptr, op->d_lp_mid, op->u_tgt, op->Mp, op->b, op->nl, m, st);
            run_sweep(op, 1, nullptr, op->d_lp_mid, op->d_target, nullptr, st);
        } else {
            launch_scale_rows(x, op->d_dc, op->d_scaled, op->cols, op->b, m, st);
            run_sweep(op, 0, op->d_scaled, nullptr, op->d_mid, nullptr, st);
            run_sweep(op, 1, op->d_mid, nullptr, op->d_target, nullptr, st);
        }
    } else {
        launch_scale_rows(x, op->d_dc, op->d_scaled, op->cols, op->b, m, st);
        launch_cyclic_fwd(op->dm, op->d_scaled, op->d_t, nullptr, op->b, m, st);
        launch_elim_y(op->dm, op->d2, op->d_zoff, op->d_zkappa, op->d_R, op->d_Bk, op->n_v, op->kdim,
                      op->d_scaled + op->x_cols * op->b, op->d_t, op->d_y, op->b, m, st);
        launch_cyclic_fwd(op->d2, op->d_y, op->d_target, nullptr, op->b, m, st);
    }
}

// out = Dc F^T (op->d_target)
void backward_from_target(pgw_op* op, uint32_t* out, hipStream_t st) {
    const ModP& m = op->m;
    if (op->kind == PGW_KIND_SYNTHETIC) {
        if (op->method == PGW_METHOD_MFMA) {
            launch_limb_split(op->d_target, nullptr, op->d_lp_src, op->u_src, op->M, op->b, op->nl, m, st);
            run_sweep(op, 2, nullptr, op->d_lp_src, op->d_mid, nullptr, st);
            launch_limb_split(op->d_mid, nullptr, op->d_lp_mid, op->u_tgt, op->Mp, op->b, op->nl, m, st);
            run_sweep(op, 3, nullptr, op->d_lp_mid, out, op->d_dc, st);
        } else {
            run_sweep(op, 2, op->d_target, nullptr, op->d_mid, nullptr, st);
            run_sweep(op, 3, op->d_mid, nullptr, out, op->d_dc, st);
        }
    } else {
        launch_cyclic_trn(op->d2, op->d_target, op->d_h, nullptr, op->b, m, st);
        launch_elim_rt_h(op->dm, op->d2, op->d_R, op->n_v, op->d_h, op->d_tadj, op->b, m, st);
        // -x_adj scaled by Dc: fold the sign into the scale (p - Dc)
        launch_cyclic_trn(op->dm, op->d_tadj, out, op->d_negdc, op->b, m, st);
        launch_elim_bt_h(op->d2, op->z_count, op->d_zrank, op->d_zkappa, op->d_Bk, op->kdim, op->d_h,
                         out + op->x_cols * op->b, op->d_dc + op->x_cols, op->b, m, st);
    }
}

void apply_square_impl(pgw_op* op, const uint32_t* in, uint32_t* out, hipStream_t st) {
    forward_into_target(op, in, st);
    launch_target_precond(op->d_target, op->d_pairs, op->pair_count, op->rounds, op->d_dr,
                          op->unpaired, op->rows, op->b, op->m, st);
    backward_from_target(op, out, st);
}

hipStream_t pick_stream(pgw_op* op, void* stream) {
    return stream ? (hipStream_t)stream : op->ctx->stream;
}

void validate_desc_common(const pgw_operator_desc* d) {
    require(d != nullptr, "null operator description");
    require(d->prime > 2 && d->prime < (1u << 31), "prime must satisfy 2 < p < 2^31");
    for (uint32_t q = 2; (uint64_t)q * q <= d->prime; ++q)
        require(d->prime % q != 0, "modulus " + std::to_string(d->prime) + " is not prime");
    require(d->block >= 1 && d->block <= 16, "block width must be in 1..16");
}

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

// ---------------------------------------------------------------------------

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
            std::strncpy(name, prop.name, name_len - 1);
            name[name_len - 1] = 0;
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
        } else if (desc->kind == PGW_KIND_CYCLIC_ELIMINATED) {
            CyclicLayout L;
            build_cyclic_layout(L, *desc);
            r = L.rows;
            c = L.cols;
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
        op->ctx = ctx;
        op->kind = desc->kind;
        op->p = desc->prime;
        op->b = desc->block;
        op->m = make_modp(op->p);
        DeviceArena& A = op->arena;
        const uint32_t b = op->b;
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
            const bool mfma_ok = mfma_block_supported(b) && op->n <= 32 && op->k >= 1;
            if (method == PGW_METHOD_AUTO) method = mfma_ok ? PGW_METHOD_MFMA : PGW_METHOD_PLAIN;
            require(method == PGW_METHOD_PLAIN || mfma_ok,
                    "MFMA path needs b in {4, 8, 16} and n <= 32", PGW_ERR_UNSUPPORTED);
            op->method = method;
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
                for (int w = 0; w < 4; ++w) {
                    DenseSweep s = sweep_desc(op, w);
                    check_mfma_bound(s, op->nl);
                    build_mfma_table(op->mt[w], op->n, s.r_out, s.k_in, op->p, src[w], !s.fwd);
                    A.ptrs.push_back(op->mt[w].d);
                    A.bytes += (uint64_t)op->n * op->nl * op->mt[w].rows_pad * op->mt[w].k_pad;
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
        } else if (desc->kind == PGW_KIND_CYCLIC_ELIMINATED) {
            require(desc->method != PGW_METHOD_MFMA,
                    "the cyclic operator's character blocks run on the plain path", PGW_ERR_UNSUPPORTED);
            op->method = PGW_METHOD_PLAIN;
            CyclicLayout L;
            build_cyclic_layout(L, *desc);
            op->rows = L.rows;
            op->cols = L.cols;
            op->mid = L.d2.col_rank.size();
            op->dm = upload_phi(A, L.dm);
            op->d2 = upload_phi(A, L.d2);
            op->d_zoff = A.upload(L.z_off);
            op->d_zrank = A.upload(L.z_rank);
            op->d_zkappa = A.upload(L.z_kappa);
            const uint32_t g = desc->genus, nv = g - 3;
            op->n_v = nv;
            op->kdim = desc->kernel_dim;
            for (uint64_t q = 0; q < (uint64_t)nv * nv * g; ++q)
                require(desc->mu[q] < op->p, "mu.values entry is not reduced modulo the instance modulus");
            op->d_R = A.upload(desc->right_inverse, (uint64_t)g * nv);
            op->d_Bk = A.upload(desc->kernel_basis, (uint64_t)g * desc->kernel_dim);
            op->x_cols = L.x_cols;
            op->z_count = L.z_count;
            op->d_scaled = A.alloc<uint32_t>(op->cols * b);
            op->d_t = A.alloc<uint32_t>(L.dm.row_rank.size() * b);
            op->d_y = A.alloc<uint32_t>(op->mid * b);
            op->d_h = A.alloc<uint32_t>(op->mid * b);
            op->d_tadj = A.alloc<uint32_t>(L.dm.row_rank.size() * b);
        } else {
            throw Error(PGW_ERR_INVALID, "unknown operator kind");
        }
        // preconditioners and the row-mixing plan
        for (uint64_t q = 0; q < op->rows; ++q)
            require(h_row_precond[q] >= 1 && h_row_precond[q] < op->p, "row preconditioner entries must be nonzero residues");
        for (uint64_t q = 0; q < op->cols; ++q)
            require(h_col_precond[q] >= 1 && h_col_precond[q] < op->p, "column preconditioner entries must be nonzero residues");
        op->d_dr = A.upload(h_row_precond, op->rows);
        op->d_dc = A.upload(h_col_precond, op->cols);
        std::vector<uint32_t> negdc(op->cols);
        for (uint64_t q = 0; q < op->cols; ++q) negdc[q] = op->p - h_col_precond[q];
        op->d_negdc = A.upload(negdc);
        op->rounds = rowmix_rounds;
        op->pair_count = op->rows / 2;
        if (rowmix_rounds > 0 && op->pair_count > 0) {
            require(h_pairs != nullptr, "row-mix plan is required when rounds > 0");
            const uint64_t total = (uint64_t)rowmix_rounds * op->pair_count;
            std::vector<uint8_t> seen(op->rows);
            for (uint32_t r = 0; r < rowmix_rounds; ++r) {
                std::fill(seen.begin(), seen.end(), 0);
                for (uint64_t q = 0; q < op->pair_count; ++q) {
                    const pgw_rowmix_pair& pr = h_pairs[(uint64_t)r * op->pair_count + q];
                    require(pr.first < op->rows && pr.second < op->rows && pr.first != pr.second,
                            "row-mix pair outside the target rows");
                    require(!seen[pr.first] && !seen[pr.second], "row-mix pairs of a round must be disjoint");
                    require(pr.a >= 1 && pr.a < op->p && pr.b >= 1 && pr.b < op->p,
                            "row-mix coefficients must be nonzero residues");
                    seen[pr.first] = seen[pr.second] = 1;
                }
            }
            op->unpaired = -1;
            if (op->rows % 2 == 1)
                for (uint64_t q = 0; q < op->rows; ++q)
                    if (!seen[q]) op->unpaired = (int64_t)q;  // left out of the last round
            op->d_pairs = A.upload(h_pairs, total);
        } else {
            op->rounds = 0;
        }
        op->d_target = A.alloc<uint32_t>(op->rows * b);
        op->d_gram = A.alloc<uint64_t>(gram_scratch_bytes(b) / sizeof(uint64_t));
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
        PGW_HIP(hipStreamSynchronize(op->ctx->stream));
        op->arena.release();
        for (auto& e : op->sweep_events) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
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
                    pick_stream(op, stream));
    });
}

int pgw_sequence_run(pgw_op* op, const uint32_t* h_v, uint64_t max_len, uint32_t* h_seq,
                     uint32_t* h_curv, uint64_t* out_len, pgw_progress_fn progress, void* user) {
    uint32_t* d_B = nullptr;
    uint32_t* d_D = nullptr;
    uint32_t* d_seq = nullptr;
    int rc = guarded([&] {
        require(op && h_v && h_seq && h_curv, "null operator or buffer");
        require(max_len >= 1, "sequence length must be positive");
        PGW_HIP(hipSetDevice(op->ctx->device));
        hipStream_t st = op->ctx->stream;
        const uint64_t N = op->cols, b = op->b, ut = b * (b + 1) / 2;
        const uint64_t rounds = (max_len - 1 + 3) / 4;
        const uint64_t L = 1 + 4 * rounds;
        for (uint64_t q = 0; q < N * b; ++q) require(h_v[q] < op->p, "V entries must be reduced modulo p");
        PGW_HIP(hipMalloc(&d_B, N * b * sizeof(uint32_t)));
        PGW_HIP(hipMalloc(&d_D, N * b * sizeof(uint32_t)));
        PGW_HIP(hipMalloc(&d_seq, L * ut * sizeof(uint32_t)));
        PGW_HIP(hipMemcpyAsync(d_B, h_v, N * b * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        uint64_t pos = 0;
        launch_gram(d_B, d_B, N, op->b, 0, d_seq + ut * pos, nullptr, op->d_gram, op->m, st);
        ++pos;
        for (uint64_t r = 0; r < rounds; ++r) {
            apply_square_impl(op, d_B, d_D, st);
            launch_gram(d_B, d_D, N, op->b, 1, d_seq + ut * pos, d_seq + ut * (pos + 1), op->d_gram, op->m, st);
            pos += 2;
            apply_square_impl(op, d_D, d_B, st);
            launch_gram(d_B, d_D, N, op->b, 2, d_seq + ut * pos, d_seq + ut * (pos + 1), op->d_gram, op->m, st);
            pos += 2;
            if (progress && ((r + 1) % 64 == 0 || r + 1 == rounds)) {
                PGW_HIP(hipStreamSynchronize(st));  // also bounds the launch queue
                progress(pos, L, user);
            }
        }
        PGW_HIP(hipMemcpyAsync(h_seq, d_seq, L * ut * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        PGW_HIP(hipMemcpyAsync(h_curv, d_B, N * b * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        PGW_HIP(hipStreamSynchronize(st));
        if (out_len) *out_len = L;
    });
    if (d_B) (void)hipFree(d_B);
    if (d_D) (void)hipFree(d_D);
    if (d_seq) (void)hipFree(d_seq);
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
        PGW_HIP(hipSetDevice(op->ctx->device));
        hipStream_t st = op->ctx->stream;
        const int count = mfma_config_count();
        if (op->method != PGW_METHOD_MFMA) {
            for (int w = 0; w < 4; ++w) {
                if (h_choice) h_choice[w] = 0;
                if (h_ms) h_ms[w] = 0.0;
            }
            return;
        }
        if (reps == 0) reps = 1;
        hipEvent_t e0, e1;
        PGW_HIP(hipEventCreate(&e0));
        PGW_HIP(hipEventCreate(&e1));
        // the sweeps' schedules do not depend on the data: time them on the
        // operator's own digit planes and scratch blocks
        for (int w = 0; w < 4; ++w) {
            const DenseSweep s = sweep_desc(op, w);
            const int8_t* lp = s.fwd ? op->d_lp_src : op->d_lp_mid;
            uint32_t* out = s.fwd ? op->d_mid : op->d_target;
            double best = 1e30;
            int best_cfg = op->mfma_cfg[w];
            for (int c = 0; c < count; ++c) {
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
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

int pgw_bench_apply(pgw_op* op, uint32_t* d_a, uint32_t* d_b, uint32_t iters, double* total_ms,
                    double* sweep_avg_ms, uint32_t* sweeps_per_apply) {
    return guarded([&] {
        require(op && d_a && d_b && iters >= 1, "null operator/block or zero iterations");
        PGW_HIP(hipSetDevice(op->ctx->device));
        hipStream_t st = op->ctx->stream;
        for (auto& e : op->sweep_events) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        op->sweep_events.clear();
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
        if (total_ms) *total_ms = ms;
        // per-launch sweep durations: one more apply with events around each sweep
        op->profile = op->kind == PGW_KIND_SYNTHETIC;
        apply_square_impl(op, (iters % 2) ? d_b : d_a, (iters % 2) ? d_a : d_b, st);
        op->profile = false;
        PGW_HIP(hipStreamSynchronize(st));
        double sum = 0;
        for (auto& e : op->sweep_events) {
            float s = 0;
            PGW_HIP(hipEventElapsedTime(&s, e.first, e.second));
            sum += s;
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        const size_t count = op->sweep_events.size();
        op->sweep_events.clear();
        if (sweep_avg_ms) *sweep_avg_ms = count ? sum / (double)count : 0.0;
        if (sweeps_per_apply) *sweeps_per_apply = (uint32_t)count;
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
    });
}

}  // extern "C"

// ===========================================================================
// batches: many sector jobs advanced together, one captured graph per round
// ===========================================================================

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

namespace {

struct BatchJob {
    pgw_op* op = nullptr;
    uint32_t* d_B = nullptr;
    uint32_t* d_D = nullptr;
    uint32_t* d_seq = nullptr;
    uint32_t* d_pos = nullptr;
    uint64_t L = 0, rounds_total = 0, rounds_done = 0;
    bool done = false;
    uint64_t bytes = 0;

    void release() {
        for (uint32_t* p : {d_B, d_D, d_seq, d_pos})
            if (p) (void)hipFree(p);
        d_B = d_D = d_seq = d_pos = nullptr;
    }
};

// One round of the sequence loop for a job (cuprym_cyclic.cu:2439-2449):
// D = S B; seq[pos] = B^T D, seq[pos+1] = D^T D; B = S D; seq[pos+2] = B^T D,
// seq[pos+3] = B^T B; pos += 4.  Positions are read on the device.
void enqueue_round(BatchJob& j, hipStream_t st) {
    pgw_op* op = j.op;
    const uint64_t ut = (uint64_t)op->b * (op->b + 1) / 2;
    apply_square_impl(op, j.d_B, j.d_D, st);
    launch_gram(j.d_B, j.d_D, op->cols, op->b, 1, j.d_seq, j.d_seq + ut, op->d_gram, op->m, st,
                j.d_pos, 0);
    apply_square_impl(op, j.d_D, j.d_B, st);
    launch_gram(j.d_B, j.d_D, op->cols, op->b, 2, j.d_seq, j.d_seq + ut, op->d_gram, op->m, st,
                j.d_pos, 2);
    launch_advance_pos(j.d_pos, 4, st);
}

}  // namespace

struct pgw_batch {
    pgw_ctx* ctx = nullptr;
    std::vector<hipStream_t> streams;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;
    std::map<uint64_t, BatchJob> jobs;
    uint64_t next_id = 1;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<uint64_t> graph_jobs;  // active set the graph was captured for
    bool use_graph = true;
    uint64_t graph_builds = 0;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_jobs.clear();
    }

    std::vector<uint64_t> active() const {
        std::vector<uint64_t> a;
        for (const auto& kv : jobs)
            if (!kv.second.done) a.push_back(kv.first);
        return a;
    }

    // enqueue one round of every active job, spread over the streams, joined on streams[0]
    void enqueue_all(const std::vector<uint64_t>& act) {
        hipStream_t s0 = streams[0];
        PGW_HIP(hipEventRecord(fork, s0));
        const size_t S = streams.size();
        for (size_t q = 1; q < S && q < act.size(); ++q) PGW_HIP(hipStreamWaitEvent(streams[q], fork, 0));
        for (size_t q = 0; q < act.size(); ++q) enqueue_round(jobs.at(act[q]), streams[q % S]);
        for (size_t q = 1; q < S && q < act.size(); ++q) {
            PGW_HIP(hipEventRecord(joins[q], streams[q]));
            PGW_HIP(hipStreamWaitEvent(s0, joins[q], 0));
        }
    }

    void step(uint32_t max_rounds) {
        std::vector<uint64_t> act = active();
        if (act.empty()) return;
        uint64_t r = max_rounds ? max_rounds : 1;
        for (uint64_t id : act) r = std::min(r, jobs.at(id).rounds_total - jobs.at(id).rounds_done);
        hipStream_t s0 = streams[0];
        if (use_graph) {
            if (act != graph_jobs) {
                drop_graph();
                PGW_HIP(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
                try {
                    enqueue_all(act);
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(s0, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                PGW_HIP(hipStreamEndCapture(s0, &graph));
                PGW_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
                graph_jobs = act;
                ++graph_builds;
            }
            for (uint64_t q = 0; q < r; ++q) PGW_HIP(hipGraphLaunch(exec, s0));
        } else {
            for (uint64_t q = 0; q < r; ++q) enqueue_all(act);
        }
        PGW_HIP(hipStreamSynchronize(s0));
        for (uint64_t id : act) {
            BatchJob& j = jobs.at(id);
            j.rounds_done += r;
            if (j.rounds_done >= j.rounds_total) j.done = true;
        }
    }
};

extern "C" {

int pgw_batch_create(pgw_ctx* ctx, uint32_t nstreams, int use_graph, pgw_batch** out) {
    pgw_batch* b = nullptr;
    int rc = guarded([&] {
        require(ctx && out, "null context or output pointer");
        PGW_HIP(hipSetDevice(ctx->device));
        b = new pgw_batch;
        b->ctx = ctx;
        b->use_graph = use_graph != 0;
        if (nstreams == 0) nstreams = 8;
        b->streams.resize(nstreams, nullptr);
        b->joins.resize(nstreams, nullptr);
        for (uint32_t q = 0; q < nstreams; ++q) {
            PGW_HIP(hipStreamCreateWithFlags(&b->streams[q], hipStreamNonBlocking));
            PGW_HIP(hipEventCreateWithFlags(&b->joins[q], hipEventDisableTiming));
        }
        PGW_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
        *out = b;
    });
    if (rc != PGW_OK && b) {
        for (auto s : b->streams) if (s) (void)hipStreamDestroy(s);
        for (auto e : b->joins) if (e) (void)hipEventDestroy(e);
        if (b->fork) (void)hipEventDestroy(b->fork);
        delete b;
    }
    return rc;
}

int pgw_batch_add(pgw_batch* b, pgw_op* op, const uint32_t* h_v, uint64_t max_len, uint64_t* job) {
    BatchJob j;
    int rc = guarded([&] {
        require(b && op && h_v && job, "null batch, operator or buffer");
        require(op->ctx == b->ctx, "operator belongs to another context");
        require(max_len >= 1, "sequence length must be positive");
        for (const auto& kv : b->jobs)
            require(kv.second.done || kv.second.op != op, "an operator can carry one active job at a time");
        PGW_HIP(hipSetDevice(b->ctx->device));
        const uint64_t N = op->cols, bb = op->b, ut = bb * (bb + 1) / 2;
        for (uint64_t q = 0; q < N * bb; ++q) require(h_v[q] < op->p, "V entries must be reduced modulo p");
        j.op = op;
        j.rounds_total = (max_len - 1 + 3) / 4;
        j.L = 1 + 4 * j.rounds_total;
        PGW_HIP(hipMalloc(&j.d_B, N * bb * sizeof(uint32_t)));
        PGW_HIP(hipMalloc(&j.d_D, N * bb * sizeof(uint32_t)));
        PGW_HIP(hipMalloc(&j.d_seq, j.L * ut * sizeof(uint32_t)));
        PGW_HIP(hipMalloc(&j.d_pos, sizeof(uint32_t)));
        j.bytes = 2 * N * bb * sizeof(uint32_t) + j.L * ut * sizeof(uint32_t);
        hipStream_t s0 = b->streams[0];
        const uint32_t one = 1;
        PGW_HIP(hipMemcpyAsync(j.d_B, h_v, N * bb * sizeof(uint32_t), hipMemcpyHostToDevice, s0));
        PGW_HIP(hipMemcpyAsync(j.d_pos, &one, sizeof(uint32_t), hipMemcpyHostToDevice, s0));
        launch_gram(j.d_B, j.d_B, N, op->b, 0, j.d_seq, nullptr, op->d_gram, op->m, s0);  // seq[0] = V^T V
        PGW_HIP(hipStreamSynchronize(s0));
        if (j.rounds_total == 0) j.done = true;
        const uint64_t id = b->next_id++;
        b->jobs.emplace(id, j);
        *job = id;
    });
    if (rc != PGW_OK) j.release();
    return rc;
}

int pgw_batch_step(pgw_batch* b, uint32_t max_rounds, uint64_t* active) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        PGW_HIP(hipSetDevice(b->ctx->device));
        b->step(max_rounds);
        if (active) *active = b->active().size();
    });
}

int pgw_batch_poll(pgw_batch* b, uint64_t job, int* done) {
    return guarded([&] {
        require(b && done, "null batch or output");
        auto it = b->jobs.find(job);
        require(it != b->jobs.end(), "unknown batch job");
        *done = it->second.done ? 1 : 0;
    });
}

int pgw_batch_result(pgw_batch* b, uint64_t job, uint32_t* h_seq, uint32_t* h_curv, uint64_t* len) {
    return guarded([&] {
        require(b && h_seq && h_curv, "null batch or buffer");
        auto it = b->jobs.find(job);
        require(it != b->jobs.end(), "unknown batch job");
        BatchJob& j = it->second;
        require(j.done, "batch job has not finished");
        PGW_HIP(hipSetDevice(b->ctx->device));
        const uint64_t N = j.op->cols, bb = j.op->b, ut = bb * (bb + 1) / 2;
        PGW_HIP(hipMemcpy(h_seq, j.d_seq, j.L * ut * sizeof(uint32_t), hipMemcpyDeviceToHost));
        PGW_HIP(hipMemcpy(h_curv, j.d_B, N * bb * sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (len) *len = j.L;
        j.release();
        b->jobs.erase(it);
    });
}

int pgw_batch_cancel(pgw_batch* b, uint64_t job) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        auto it = b->jobs.find(job);
        require(it != b->jobs.end(), "unknown batch job");
        PGW_HIP(hipSetDevice(b->ctx->device));
        for (auto s : b->streams) PGW_HIP(hipStreamSynchronize(s));
        b->drop_graph();
        it->second.release();
        b->jobs.erase(it);
    });
}

int pgw_batch_destroy(pgw_batch* b) {
    return guarded([&] {
        if (!b) return;
        PGW_HIP(hipSetDevice(b->ctx->device));
        for (auto s : b->streams) (void)hipStreamSynchronize(s);
        b->drop_graph();
        for (auto& kv : b->jobs) kv.second.release();
        for (auto s : b->streams) (void)hipStreamDestroy(s);
        for (auto e : b->joins) (void)hipEventDestroy(e);
        (void)hipEventDestroy(b->fork);
        delete b;
    });
}

int pgw_batch_graph_builds(const pgw_batch* b, uint64_t* builds) {
    return guarded([&] {
        require(b && builds, "null batch or output");
        *builds = b->graph_builds;
    });
}

}  // extern "C"

// ===========================================================================
// serving: a device-resident loop that admits sector jobs by HBM footprint
// ===========================================================================

namespace {

using Clock = std::chrono::steady_clock;

struct OwnedDesc {
    pgw_operator_desc d{};
    std::vector<uint32_t> table_d, table_e, mu, right_inverse, kernel_basis;
    std::vector<uint16_t> wv, wa0, wa1, kw;

    void copy_from(const pgw_operator_desc& s) {
        d = s;
        if (s.kind == PGW_KIND_SYNTHETIC) {
            const uint64_t cnt = (uint64_t)s.n * s.m_tgt * s.m_src;
            table_d.assign(s.table_d, s.table_d + cnt);
            table_e.assign(s.table_e, s.table_e + cnt);
            d.table_d = table_d.data();
            d.table_e = table_e.data();
        } else {
            const uint64_t n = s.genus - 3;
            auto grab = [](auto& dst, const auto* src, uint64_t cnt) {
                if (src && cnt) dst.assign(src, src + cnt);
            };
            grab(mu, s.mu, n * n * s.genus);
            grab(wv, s.weight_v, n);
            grab(wa0, s.weight_a0, s.genus);
            grab(wa1, s.weight_a1, n);
            grab(kw, s.kernel_weights, s.kernel_dim);
            grab(right_inverse, s.right_inverse, (uint64_t)s.genus * n);
            grab(kernel_basis, s.kernel_basis, (uint64_t)s.genus * s.kernel_dim);
            d.mu = mu.data();
            d.weight_v = wv.data();
            d.weight_a0 = wa0.data();
            d.weight_a1 = wa1.data();
            d.kernel_weights = kw.data();
            d.right_inverse = right_inverse.data();
            d.kernel_basis = kernel_basis.data();
        }
    }
};

struct ServerJob {
    uint64_t ticket = 0;
    OwnedDesc desc;
    std::vector<uint32_t> dr, dc, v;
    std::vector<pgw_rowmix_pair> pairs;
    uint32_t rounds = 0;
    uint64_t max_len = 0, rows = 0, cols = 0, estimate = 0;
    pgw_op* op = nullptr;
    uint64_t batch_job = 0;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    std::string error;
    std::vector<uint32_t> seq, curv;
    uint64_t L = 0;
    Clock::time_point t_submit, t_admit, t_finish;
};

// device bytes a job needs: operator (tables, digit planes, scratch blocks) plus
// its sequence state (two N x b blocks and L x b(b+1)/2 terms)
uint64_t estimate_job_bytes(const pgw_operator_desc& d, uint64_t rows, uint64_t cols,
                            uint64_t max_len, uint32_t rowmix_rounds) {
    const uint64_t b = d.block, ut = b * (b + 1) / 2;
    const uint64_t L = 1 + 4 * ((max_len - 1 + 3) / 4);
    uint64_t bytes = 2 * cols * b * 4 + L * ut * 4;            // job state
    bytes += (rows + 2 * cols) * 4 + rows * b * 4;             // Dr, Dc, -Dc, target block
    bytes += gram_scratch_bytes((uint32_t)b) + rows / 2 * 16 * rowmix_rounds;  // Gram scratch, pairs
    if (d.kind == PGW_KIND_SYNTHETIC) {
        const uint64_t mid = host_binom(d.n, d.k - 1) * d.m_tgt;
        const uint64_t us = host_binom(d.n, d.k), ut2 = host_binom(d.n, d.k - 1);
        bytes += mid * b * 4 + cols * b * 4;                   // mid block, scaled block
        const uint64_t rp = (std::max(d.m_src, d.m_tgt) + 15) / 16 * 16 + 16;
        const uint64_t kp = (std::max(d.m_src, d.m_tgt) + 31) / 32 * 32 + 32;
        bytes += 4 * (uint64_t)d.n * 4 * rp * kp * 2;          // tables (either path, upper bound)
        bytes += 2 * us * 4 * b * (((uint64_t)d.m_src + 15) / 16 * 16);   // source digit planes
        bytes += 2 * ut2 * 4 * b * (((uint64_t)d.m_tgt + 15) / 16 * 16);  // mid digit planes
    } else {
        bytes += 6 * (rows + cols) * b * 4 + 64ull * (rows + cols);
    }
    return bytes + (16ull << 20);
}

}  // namespace

struct pgw_server {
    pgw_ctx* ctx = nullptr;
    uint64_t budget = 0;
    uint32_t max_active = 0, rounds_per_step = 1;
    pgw_batch* batch = nullptr;
    std::mutex mtx;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> queue;
    std::map<uint64_t, std::unique_ptr<ServerJob>> jobs;
    std::vector<uint64_t> running;
    uint64_t next_ticket = 1, in_use = 0, peak_bytes = 0, peak_active = 0, completed = 0;
    bool stop = false;
    std::string fatal;
    std::thread worker;
    Clock::time_point t0 = Clock::now();

    void finish_job(ServerJob& j, int state, const std::string& err) {
        j.state = state;
        j.error = err;
        j.t_finish = Clock::now();
        if (j.op) {
            pgw_operator_destroy(j.op);
            j.op = nullptr;
        }
    }

    void run() {
        (void)hipSetDevice(ctx->device);
        for (;;) {
            // admission: FIFO while the HBM budget and the slot count allow
            std::vector<ServerJob*> admit;
            {
                std::unique_lock<std::mutex> lk(mtx);
                if (running.empty())
                    cv_work.wait(lk, [&] { return stop || !queue.empty(); });
                if (stop && queue.empty() && running.empty()) return;
                while (!queue.empty() && running.size() + admit.size() < max_active) {
                    ServerJob& j = *jobs.at(queue.front());
                    if (j.estimate > budget) {
                        queue.pop_front();
                        finish_job(j, 3, "job needs more device memory than the server budget");
                        cv_done.notify_all();
                        continue;
                    }
                    if (in_use + j.estimate > budget) break;
                    in_use += j.estimate;
                    peak_bytes = std::max(peak_bytes, in_use);
                    queue.pop_front();
                    admit.push_back(&j);
                }
            }
            for (ServerJob* j : admit) {
                j->t_admit = Clock::now();
                int rc = pgw_operator_create(ctx, &j->desc.d, j->dr.data(), j->dc.data(), j->rounds,
                                             j->pairs.empty() ? nullptr : j->pairs.data(), &j->op);
                if (rc == PGW_OK)
                    rc = pgw_batch_add(batch, j->op, j->v.data(), j->max_len, &j->batch_job);
                std::lock_guard<std::mutex> lk(mtx);
                if (rc != PGW_OK) {
                    in_use -= j->estimate;
                    finish_job(*j, 3, pgw_last_error());
                    cv_done.notify_all();
                } else {
                    // replace the estimate by what the operator and the job really hold
                    pgw_op_info info{};
                    pgw_operator_info(j->op, &info);
                    const uint64_t b = j->desc.d.block, ut = b * (b + 1) / 2;
                    const uint64_t L = 1 + 4 * ((j->max_len - 1 + 3) / 4);
                    const uint64_t actual = info.device_bytes + 2 * j->cols * b * 4 + L * ut * 4;
                    in_use = in_use - j->estimate + actual;
                    j->estimate = actual;
                    peak_bytes = std::max(peak_bytes, in_use);
                    running.push_back(j->ticket);
                    peak_active = std::max<uint64_t>(peak_active, running.size());
                }
                j->dr = std::vector<uint32_t>();
                j->dc = std::vector<uint32_t>();
                j->v = std::vector<uint32_t>();
                j->pairs = std::vector<pgw_rowmix_pair>();
            }
            if (running.empty()) continue;
            uint64_t active = 0;
            if (pgw_batch_step(batch, rounds_per_step, &active) != PGW_OK) {
                std::lock_guard<std::mutex> lk(mtx);
                fatal = pgw_last_error();
                for (uint64_t t : running) {
                    pgw_batch_cancel(batch, jobs.at(t)->batch_job);
                    finish_job(*jobs.at(t), 3, fatal);
                    in_use -= jobs.at(t)->estimate;
                }
                running.clear();
                cv_done.notify_all();
                continue;
            }
            // completions: stream results to host memory, free the device state
            std::vector<uint64_t> still;
            for (uint64_t t : running) {
                ServerJob& j = *jobs.at(t);
                int done = 0;
                pgw_batch_poll(batch, j.batch_job, &done);
                if (!done) {
                    still.push_back(t);
                    continue;
                }
                const uint64_t b = j.desc.d.block, ut = b * (b + 1) / 2;
                const uint64_t L = 1 + 4 * ((j.max_len - 1 + 3) / 4);
                j.seq.resize(L * ut);
                j.curv.resize(j.cols * b);
                int rc = pgw_batch_result(batch, j.batch_job, j.seq.data(), j.curv.data(), &j.L);
                std::lock_guard<std::mutex> lk(mtx);
                in_use -= j.estimate;
                finish_job(j, rc == PGW_OK ? 2 : 3, rc == PGW_OK ? "" : pgw_last_error());
                ++completed;
                cv_done.notify_all();
            }
            std::lock_guard<std::mutex> lk(mtx);
            running.swap(still);
        }
    }
};

extern "C" {

int pgw_job_footprint(const pgw_operator_desc* desc, uint64_t max_len, uint32_t rowmix_rounds,
                      uint64_t* bytes) {
    return guarded([&] {
        require(desc && bytes, "null description or output");
        require(max_len >= 1, "sequence length must be positive");
        uint64_t rows = 0, cols = 0;
        if (pgw_operator_dims(desc, &rows, &cols) != PGW_OK) throw Error(PGW_ERR_INVALID, pgw_last_error());
        *bytes = estimate_job_bytes(*desc, rows, cols, max_len, rowmix_rounds);
    });
}

int pgw_server_create(pgw_ctx* ctx, uint64_t hbm_budget, uint32_t max_active,
                      uint32_t rounds_per_step, pgw_server** out) {
    pgw_server* s = nullptr;
    int rc = guarded([&] {
        require(ctx && out, "null context or output pointer");
        require(hbm_budget > 0 && max_active > 0, "budget and slot count must be positive");
        s = new pgw_server;
        s->ctx = ctx;
        s->budget = hbm_budget;
        s->max_active = max_active;
        s->rounds_per_step = rounds_per_step ? rounds_per_step : 1;
        int brc = pgw_batch_create(ctx, std::min<uint32_t>(max_active, 8), 1, &s->batch);
        if (brc != PGW_OK) throw Error(brc, pgw_last_error());
        s->worker = std::thread([s] { s->run(); });
        *out = s;
    });
    if (rc != PGW_OK && s) {
        if (s->batch) pgw_batch_destroy(s->batch);
        delete s;
    }
    return rc;
}

int pgw_server_submit(pgw_server* s, const pgw_operator_desc* desc, const uint32_t* h_dr,
                      const uint32_t* h_dc, uint32_t rowmix_rounds, const pgw_rowmix_pair* h_pairs,
                      const uint32_t* h_v, uint64_t max_len, uint64_t* ticket) {
    return guarded([&] {
        require(s && desc && h_dr && h_dc && h_v && ticket, "null server, description or buffer");
        require(max_len >= 1, "sequence length must be positive");
        uint64_t rows = 0, cols = 0;
        if (pgw_operator_dims(desc, &rows, &cols) != PGW_OK) throw Error(PGW_ERR_INVALID, pgw_last_error());
        auto j = std::make_unique<ServerJob>();
        j->t_submit = Clock::now();
        j->desc.copy_from(*desc);
        j->dr.assign(h_dr, h_dr + rows);
        j->dc.assign(h_dc, h_dc + cols);
        j->v.assign(h_v, h_v + cols * desc->block);
        j->rounds = rowmix_rounds;
        if (rowmix_rounds && rows / 2) {
            require(h_pairs != nullptr, "row-mix plan is required when rounds > 0");
            j->pairs.assign(h_pairs, h_pairs + (uint64_t)rowmix_rounds * (rows / 2));
        }
        j->max_len = max_len;
        j->rows = rows;
        j->cols = cols;
        j->estimate = estimate_job_bytes(*desc, rows, cols, max_len, rowmix_rounds);
        std::lock_guard<std::mutex> lk(s->mtx);
        require(!s->stop, "server is shutting down");
        j->ticket = s->next_ticket++;
        *ticket = j->ticket;
        s->queue.push_back(j->ticket);
        s->jobs.emplace(j->ticket, std::move(j));
        s->cv_work.notify_one();
    });
}

int pgw_server_wait(pgw_server* s, uint64_t ticket, double timeout_s, int* state) {
    return guarded([&] {
        require(s && state, "null server or output");
        std::unique_lock<std::mutex> lk(s->mtx);
        auto it = s->jobs.find(ticket);
        require(it != s->jobs.end(), "unknown ticket");
        ServerJob& j = *it->second;
        auto pred = [&] { return j.state >= 2; };
        if (timeout_s < 0) s->cv_done.wait(lk, pred);
        else s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
        *state = j.state;
    });
}

int pgw_server_result(pgw_server* s, uint64_t ticket, uint32_t* h_seq, uint32_t* h_curv,
                      uint64_t* len, double* times3) {
    return guarded([&] {
        require(s, "null server");
        std::unique_lock<std::mutex> lk(s->mtx);
        auto it = s->jobs.find(ticket);
        require(it != s->jobs.end(), "unknown ticket");
        ServerJob& j = *it->second;
        require(j.state >= 2, "job has not finished");
        if (j.state == 3) throw Error(PGW_ERR_INVALID, "job failed: " + j.error);
        if (h_seq) std::copy(j.seq.begin(), j.seq.end(), h_seq);
        if (h_curv) std::copy(j.curv.begin(), j.curv.end(), h_curv);
        if (len) *len = j.L;
        if (times3) {
            times3[0] = std::chrono::duration<double>(j.t_submit - s->t0).count();
            times3[1] = std::chrono::duration<double>(j.t_admit - s->t0).count();
            times3[2] = std::chrono::duration<double>(j.t_finish - s->t0).count();
        }
        s->jobs.erase(it);
    });
}

int pgw_server_stats(pgw_server* s, uint64_t* completed, uint64_t* peak_bytes,
                     uint64_t* peak_active, uint64_t* queued) {
    return guarded([&] {
        require(s != nullptr, "null server");
        std::lock_guard<std::mutex> lk(s->mtx);
        if (completed) *completed = s->completed;
        if (peak_bytes) *peak_bytes = s->peak_bytes;
        if (peak_active) *peak_active = s->peak_active;
        if (queued) *queued = s->queue.size();
    });
}

int pgw_server_destroy(pgw_server* s) {
    return guarded([&] {
        if (!s) return;
        {
            std::lock_guard<std::mutex> lk(s->mtx);
            s->stop = true;
        }
        s->cv_work.notify_all();
        if (s->worker.joinable()) s->worker.join();
        for (auto& kv : s->jobs)
            if (kv.second->op) pgw_operator_destroy(kv.second->op);
        pgw_batch_destroy(s->batch);
        delete s;
    });
}

}  // extern "C"

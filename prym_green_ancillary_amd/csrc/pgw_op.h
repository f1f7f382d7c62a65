// pgw_op.h -- internals shared by the C ABI translation units: error
// guarding, the device-memory arena and the operator object.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "pgw_kernels.h"

namespace pgw {

// Runs f, turning exceptions into a status code and the thread's last error.
template <typename F>
inline int guarded(F&& f) {
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

// Device allocations of one operator.  Each HIP call holds `lock` (the
// context's capture mutex) when one is set, so that another thread's graph
// capture never sees an allocation, copy or free in the middle.
struct DeviceArena {
    std::vector<void*> ptrs;
    uint64_t bytes = 0;
    std::mutex* lock = nullptr;

    template <typename T>
    T* alloc(uint64_t count) {
        if (count == 0) count = 1;
        void* p = nullptr;
        hipError_t e;
        {
            std::unique_lock<std::mutex> g = guard();
            e = hipMalloc(&p, count * sizeof(T));
        }
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
        if (count) {
            std::unique_lock<std::mutex> g = guard();
            PGW_HIP(hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice));
        }
        return d;
    }
    template <typename T>
    T* upload(const std::vector<T>& v) { return upload(v.data(), (uint64_t)v.size()); }
    void release() {
        std::unique_lock<std::mutex> g = guard();
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
        bytes = 0;
    }
    std::unique_lock<std::mutex> guard() {
        return lock ? std::unique_lock<std::mutex>(*lock) : std::unique_lock<std::mutex>();
    }
};

}  // namespace pgw

// ---------------------------------------------------------------------------

struct pgw_op {
    using ModP = pgw::ModP;
    using DeviceArena = pgw::DeviceArena;
    using MfmaTable = pgw::MfmaTable;
    using SellDevice = pgw::SellDevice;
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
    // synthetic operators: the row-mixing chains (n = 0: the pair passes)
    // and the block P^T Dr P writes
    pgw::MixChainsDevice chains;
    uint32_t* d_target2 = nullptr;
    uint64_t* d_gram = nullptr;
    // block projection on int8 MFMA (else on the vector ALUs): by default for
    // operators on the MFMA path, pgw_operator_set_gram to change it
    bool gram_mfma = false;
    // synthetic
    uint32_t n = 0, k = 0, M = 0, Mp = 0;
    uint64_t u_src = 0, u_tgt = 0;
    uint32_t* d_tab[4][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    uint32_t k4[4] = {0, 0, 0, 0};
    MfmaTable mt[4];
    int mfma_cfg[4] = {8, 8, 8, 8};  // tile configuration per sweep (pgw_operator_autotune)
    uint32_t nl = 0;
    int8_t* d_lp_src = nullptr;
    int8_t* d_lp_mid = nullptr;
    uint32_t* d_inc[2] = {nullptr, nullptr};  // incidence tables: forward, transpose sweeps
    // cyclic eliminated and its deformation variants: S = M1^T M2^T M3 M2 M1
    // (explicit sparse maps, pgw_sparse.cpp); d_y holds y' (M1's rows), d_t2
    // the mixed target block
    SellDevice m1, m2, m3, m2t, m1t;
    // t in the row-mixing order (mix_order): d_t_pos[row] = its position,
    // d_t_row[position] = its row; m3 is then stored in that order.  Null:
    // t and m3 in natural row order.
    const uint32_t* d_t_pos = nullptr;
    const uint32_t* d_t_row = nullptr;
    uint64_t y_rows = 0;
    uint32_t* d_y = nullptr;
    uint32_t* d_t2 = nullptr;
    // per-stage timing of the apply chain (pgw_bench_apply, pgw_bench_stages):
    // while `profile` is set every stage launch is bracketed by two events
    struct StageEvent {
        int stage;
        double bytes;  // algorithmic bytes of the launch
        hipEvent_t e0, e1;
    };
    bool profile = false;
    std::vector<StageEvent> stage_events;
};

namespace pgw {

// d_out = S d_in on stream st (the block-apply chain, enqueue only).
void apply_square_impl(pgw_op* op, const uint32_t* in, uint32_t* out, hipStream_t st);
// Half a sequence round: out = S in, then the mode's two projections into
// seq0 / seq1 at row *pos + pos_off (mode 1: in^T out and out^T out; mode 2:
// out^T in and out^T out), the same values as apply_square_impl followed by
// launch_gram(x, y, ..., mode) with (x, y) = (in, out) or (out, in).  Cyclic
// operators at b = 4, p < 2^16 fold the projection into their last map.
void apply_and_project(pgw_op* op, const uint32_t* in, uint32_t* out, int mode, uint32_t* seq0,
                       uint32_t* seq1, const uint32_t* pos, uint32_t pos_off, hipStream_t st);

// Times tile configurations cfg_lo..cfg_hi of each of the four MFMA sweeps on
// the operator's own buffers (reps launches each, HIP events on the
// context's stream) and keeps the fastest per sweep in op->mfma_cfg; plain
// operators report configuration 0 and 0 ms.
void autotune_sweeps(pgw_op* op, uint32_t reps, int cfg_lo, int cfg_hi, uint32_t* h_choice, double* h_ms);

}  // namespace pgw

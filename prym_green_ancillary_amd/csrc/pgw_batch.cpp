// pgw_batch.cpp -- batches of sector sequence loops advanced by replayed HIP
// graphs, and the serving loop that feeds them.
//
// The reference runs one sector per process, one after another
// (scripts/reproduce.py:235-276 around cuda/cuprym_cyclic.cu:2316-2461).  Here
// one process keeps many sectors resident in HBM and advances them together:
// small sectors cannot fill 256 CUs on their own, several of them can.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <string>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "pgw_op.h"

using namespace pgw;

// ===========================================================================
// batches
// ===========================================================================
//
// Every replay advances each unfinished job by one round on ONE stream.  Jobs
// of the synthetic operator on the MFMA path whose block width, limb count,
// modulus class, row-mixing rounds and tile configurations agree form a
// group: each stage of the round (digit split, the four sweeps, row mixing,
// Gram projections, position counters) is one launch for the whole group,
// over a grid in which job j owns a contiguous block range.  Other jobs run
// their own launches after the groups'.  With PGW_BATCH_GRAPH a round is
// captured once into a single-stream graph and replayed while the set of
// jobs is unchanged.

namespace {

constexpr unsigned kUploadStreams = 4;  // concurrent pgw_batch_add calls staging V

struct BatchJob {
    uint64_t id = 0;
    pgw_op* op = nullptr;
    uint32_t* d_B = nullptr;
    uint32_t* d_D = nullptr;
    uint32_t* d_seq = nullptr;
    uint32_t* d_pos = nullptr;
    uint64_t L = 0, rounds_total = 0, rounds_done = 0;
    bool done = false;

    void release() {
        for (uint32_t* p : {d_B, d_D, d_seq, d_pos})
            if (p) (void)hipFree(p);
        d_B = d_D = d_seq = d_pos = nullptr;
    }
};

// One round of the sequence loop for a job on its own (cuprym_cyclic.cu:
// 2439-2449): D = S B; seq[pos] = B^T D, seq[pos+1] = D^T D; B = S D;
// seq[pos+2] = B^T D, seq[pos+3] = B^T B; pos += 4.  Positions are read on the
// device, so the same launches (or graph) serve every round.
void enqueue_round(const BatchJob& j, hipStream_t st) {
    pgw_op* op = j.op;
    const uint64_t ut = (uint64_t)op->b * (op->b + 1) / 2;
    apply_and_project(op, j.d_B, j.d_D, 1, j.d_seq, j.d_seq + ut, j.d_pos, 0, st);
    apply_and_project(op, j.d_D, j.d_B, 2, j.d_seq, j.d_seq + ut, j.d_pos, 2, st);
    launch_advance_pos(j.d_pos, 4, st);
}

// grouping key of a job, or an empty string when it runs on its own
std::string group_key(const pgw_op* op) {
    if (op->kind != PGW_KIND_SYNTHETIC || op->method != PGW_METHOD_MFMA) return "";
    for (int w = 0; w < 4; ++w)
        if (op->mfma_cfg[w] < 7 || op->mfma_cfg[w] > 9) return "";
    std::string k = std::to_string(op->b) + "/" + std::to_string(op->nl) + "/" +
                    (op->p == 0x7fffffffu ? "m31" : "bar") + "/" + std::to_string(op->rounds);
    for (int w = 0; w < 4; ++w) k += "/" + std::to_string(op->mfma_cfg[w]);
    return k;
}

// A device array of per-job launch records and their block offsets.
template <typename T>
struct GroupArrays {
    T* d_jobs = nullptr;
    uint32_t* d_first = nullptr;
    uint64_t blocks = 0;
};

// The stage launches of one group.
struct Group {
    uint32_t b = 0, nl = 0, rounds = 0;
    bool m31 = false;
    bool persistent = false;  // PGW_BATCH_PERSISTENT: the sweeps as persistent launches
    int cfg[4] = {8, 8, 8, 8};
    std::vector<uint64_t> ids;
    std::vector<void*> allocs;
    GroupArrays<LimbJob> split_B, split_D, split_mid, split_tgt;
    GroupArrays<MfmaSweepArgs> sweep[3], sweep3_D, sweep3_B;
    GroupArrays<MixJob> mix;
    GroupArrays<GramJob> gram;
    uint32_t* d_gram_fin = nullptr;  // finish-pass block offsets (P per job and product)
    uint64_t gram_fin_blocks = 0;
    uint32_t** d_pos = nullptr;

    void release() {
        for (void* p : allocs) (void)hipFree(p);
        allocs.clear();
    }

    template <typename T>
    T* upload(const std::vector<T>& v) {
        void* d = nullptr;
        PGW_HIP(hipMalloc(&d, std::max<size_t>(1, v.size()) * sizeof(T)));
        allocs.push_back(d);
        if (!v.empty()) PGW_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return static_cast<T*>(d);
    }

    // records and block counts -> device arrays with offsets
    template <typename T>
    GroupArrays<T> make(const std::vector<T>& recs, const std::vector<uint64_t>& blocks) {
        GroupArrays<T> g;
        std::vector<uint32_t> first(recs.size() + 1, 0);
        for (size_t q = 0; q < recs.size(); ++q) {
            g.blocks += blocks[q];
            require(g.blocks < (1ULL << 31), "grouped launch grid too large");
            first[q + 1] = (uint32_t)g.blocks;
        }
        g.d_jobs = upload(recs);
        g.d_first = upload(first);
        return g;
    }

    void build(const std::vector<const BatchJob*>& js) {
        const pgw_op* op0 = js[0]->op;
        b = op0->b;
        nl = op0->nl;
        rounds = op0->rounds;
        m31 = op0->p == 0x7fffffffu;
        for (int w = 0; w < 4; ++w) cfg[w] = op0->mfma_cfg[w];
        std::vector<LimbJob> sB, sD, smid, stgt;
        std::vector<uint64_t> nsrc, nmid;
        std::vector<MfmaSweepArgs> sw[3], s3D, s3B;
        std::vector<uint64_t> nsw[3], ns3;
        std::vector<MixJob> mx;
        std::vector<uint64_t> nmx;
        std::vector<GramJob> gr;
        std::vector<uint64_t> ngr;
        std::vector<uint32_t> fin(js.size() + 1, 0);
        std::vector<uint32_t*> pos;
        const uint32_t P = b * (b + 1) / 2;
        for (const BatchJob* j : js) {
            pgw_op* op = j->op;
            ids.push_back(j->id);
            const uint32_t ks = limb_stride(op->M), kt = limb_stride(op->Mp);
            sB.push_back({j->d_B, op->d_dc, op->d_lp_src, op->u_src, op->M, ks, op->m});
            sD.push_back({j->d_D, op->d_dc, op->d_lp_src, op->u_src, op->M, ks, op->m});
            smid.push_back({op->d_mid, nullptr, op->d_lp_mid, op->u_tgt, op->Mp, kt, op->m});
            stgt.push_back({op->d_target, nullptr, op->d_lp_src, op->u_src, op->M, ks, op->m});
            nsrc.push_back(limb_group_blocks(sB.back(), b));
            nmid.push_back(limb_group_blocks(smid.back(), b));
            const int8_t* in_lp[3] = {op->d_lp_src, op->d_lp_mid, op->d_lp_src};
            uint32_t* outs[3] = {op->d_mid, op->d_target, op->d_mid};
            for (int w = 0; w < 3; ++w) {
                MfmaSweepArgs a{};
                const uint64_t blk = mfma_group_prepare(cfg[w], b, nl, sweep_desc_of(op, w), op->mt[w],
                                                        in_lp[w], outs[w], nullptr, op->m, a);
                sw[w].push_back(a);
                nsw[w].push_back(blk);
            }
            MfmaSweepArgs a3{};
            ns3.push_back(mfma_group_prepare(cfg[3], b, nl, sweep_desc_of(op, 3), op->mt[3], op->d_lp_mid,
                                             j->d_D, op->d_dc, op->m, a3));
            s3D.push_back(a3);
            a3.out = j->d_B;
            s3B.push_back(a3);
            MixJob mj{op->d_pairs, op->pair_count, op->d_target, op->d_dr, op->unpaired, op->m};
            mx.push_back(mj);
            nmx.push_back(mix_group_blocks(mj, b));
            const uint32_t gb = gram_group_blocks(op->cols);
            gr.push_back({j->d_B, j->d_D, op->cols, op->d_gram, gb, j->d_seq, j->d_seq + P, j->d_pos, op->m});
            ngr.push_back(gb);
            pos.push_back(j->d_pos);
        }
        split_B = make(sB, nsrc);
        split_D = make(sD, nsrc);
        split_mid = make(smid, nmid);
        split_tgt = make(stgt, nsrc);
        for (int w = 0; w < 3; ++w) sweep[w] = make(sw[w], nsw[w]);
        sweep3_D = make(s3D, ns3);
        sweep3_B = make(s3B, ns3);
        mix = make(mx, nmx);
        gram = make(gr, ngr);
        for (size_t q = 0; q < js.size(); ++q) fin[q + 1] = fin[q] + 2 * P;  // modes 1 and 2: two products
        gram_fin_blocks = fin.back();
        d_gram_fin = upload(fin);
        d_pos = upload(pos);
    }

    static DenseSweep sweep_desc_of(const pgw_op* op, int which) {
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

    // one block-apply of every job: in -> out (B -> D or D -> B)
    void apply(bool from_B, hipStream_t st) const {
        const uint32_t n = (uint32_t)ids.size();
        const GroupArrays<LimbJob>& src = from_B ? split_B : split_D;
        launch_limb_split_group(src.d_jobs, src.d_first, n, src.blocks, b, nl, st);
        mfma_group_launch(cfg[0], b, nl, true, sweep[0].d_jobs, sweep[0].d_first, n, sweep[0].blocks, persistent, st);
        launch_limb_split_group(split_mid.d_jobs, split_mid.d_first, n, split_mid.blocks, b, nl, st);
        mfma_group_launch(cfg[1], b, nl, false, sweep[1].d_jobs, sweep[1].d_first, n, sweep[1].blocks, persistent, st);
        // P^T Dr P: rounds 0..R-2 forward, the fused last round, R-2..0 transposed
        for (uint32_t r = 0; r + 1 < rounds; ++r)
            launch_rowmix_group(mix.d_jobs, mix.d_first, n, mix.blocks, b, 0, m31, r, st);
        launch_rowmix_group(mix.d_jobs, mix.d_first, n, mix.blocks, b, 2, m31, rounds - 1, st);
        for (uint32_t r = rounds - 1; r-- > 0;)
            launch_rowmix_group(mix.d_jobs, mix.d_first, n, mix.blocks, b, 1, m31, r, st);
        launch_limb_split_group(split_tgt.d_jobs, split_tgt.d_first, n, split_tgt.blocks, b, nl, st);
        mfma_group_launch(cfg[2], b, nl, true, sweep[2].d_jobs, sweep[2].d_first, n, sweep[2].blocks, persistent, st);
        launch_limb_split_group(split_mid.d_jobs, split_mid.d_first, n, split_mid.blocks, b, nl, st);
        const GroupArrays<MfmaSweepArgs>& s3 = from_B ? sweep3_D : sweep3_B;
        mfma_group_launch(cfg[3], b, nl, false, s3.d_jobs, s3.d_first, n, s3.blocks, persistent, st);
    }

    void round(hipStream_t st) const {
        const uint32_t n = (uint32_t)ids.size();
        apply(true, st);
        launch_gram_group(gram.d_jobs, gram.d_first, d_gram_fin, n, gram.blocks, gram_fin_blocks, b, 1, 0, st);
        apply(false, st);
        launch_gram_group(gram.d_jobs, gram.d_first, d_gram_fin, n, gram.blocks, gram_fin_blocks, b, 2, 2, st);
        launch_advance_pos_group(d_pos, n, 4, st);
    }
};

}  // namespace

struct pgw_batch {
    pgw_ctx* ctx = nullptr;
    uint32_t flags = 0;
    hipStream_t stream = nullptr;  // every replay runs here
    hipStream_t upload[kUploadStreams] = {};
    std::mutex upload_mtx[kUploadStreams];
    std::atomic<unsigned> next_upload{0};
    // `mtx` guards the job table; `step_mtx` serialises steps with result,
    // cancel and destroy
    std::mutex mtx, step_mtx;
    std::map<uint64_t, BatchJob> jobs;
    uint64_t next_id = 1;
    // PGW_BATCH_AUTOTUNE: tile configurations (7-9) per operator shape, timed
    // on the first operator of that shape and reused for the others
    std::mutex tune_mtx;
    std::map<std::string, std::array<int, 4>> tuned;
    // the plan of the current set of unfinished jobs (under step_mtx)
    std::vector<uint64_t> plan_key;
    std::vector<Group> groups;
    std::vector<const BatchJob*> solo;  // map nodes stay put until result/cancel drop the plan
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    uint64_t graph_builds = 0;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }

    // under step_mtx, with nothing of the plan in flight
    void drop_plan() {
        drop_graph();
        std::lock_guard<std::mutex> cl(ctx->capture_mtx);
        for (Group& g : groups) g.release();
        groups.clear();
        solo.clear();
        plan_key.clear();
    }

    bool plan_uses(uint64_t id) const {
        return std::find(plan_key.begin(), plan_key.end(), id) != plan_key.end();
    }

    void build_plan(const std::vector<uint64_t>& key, const std::vector<const BatchJob*>& js) {
        drop_plan();
        std::map<std::string, std::vector<const BatchJob*>> classes;
        for (const BatchJob* j : js) {
            const std::string k = (flags & PGW_BATCH_GROUPED) ? group_key(j->op) : "";
            if (k.empty() || j->op->rounds == 0 || j->op->pair_count == 0) solo.push_back(j);
            else classes[k].push_back(j);
        }
        std::lock_guard<std::mutex> cl(ctx->capture_mtx);
        for (auto& kv : classes) {
            groups.emplace_back();
            groups.back().persistent = (flags & PGW_BATCH_PERSISTENT) != 0;
            try {
                groups.back().build(kv.second);
            } catch (...) {
                groups.back().release();
                groups.pop_back();
                throw;
            }
        }
        plan_key = key;
    }

    void enqueue_all() {
        for (const Group& g : groups) g.round(stream);
        for (const BatchJob* j : solo) enqueue_round(*j, stream);
    }

    uint64_t step(uint32_t max_replays) {
        std::lock_guard<std::mutex> sl(step_mtx);
        std::vector<uint64_t> key;
        std::vector<const BatchJob*> js;
        uint64_t reps = std::max<uint32_t>(1, max_replays);
        {
            std::lock_guard<std::mutex> lk(mtx);
            for (auto& kv : jobs)
                if (!kv.second.done) {
                    key.push_back(kv.first);
                    js.push_back(&kv.second);
                    reps = std::min<uint64_t>(reps, kv.second.rounds_total - kv.second.rounds_done);
                }
        }
        if (js.empty()) return 0;
        if (key != plan_key) build_plan(key, js);
        if (flags & PGW_BATCH_GRAPH) {
            if (!exec) {
                std::lock_guard<std::mutex> cl(ctx->capture_mtx);
                PGW_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                try {
                    enqueue_all();
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(stream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                PGW_HIP(hipStreamEndCapture(stream, &graph));
                PGW_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
                ++graph_builds;
            }
            for (uint64_t r = 0; r < reps; ++r) PGW_HIP(hipGraphLaunch(exec, stream));
        } else {
            for (uint64_t r = 0; r < reps; ++r) enqueue_all();
        }
        PGW_HIP(hipStreamSynchronize(stream));
        std::lock_guard<std::mutex> lk(mtx);
        uint64_t active = 0;
        for (uint64_t id : key) {
            BatchJob& j = jobs.at(id);
            j.rounds_done += reps;
            if (j.rounds_done >= j.rounds_total) j.done = true;
        }
        for (auto& kv : jobs) active += kv.second.done ? 0 : 1;
        return active;
    }
};

namespace {

// The operator's shape as far as the sweeps' schedules go: subsets, panel
// heights, block width, limb count and fold path.
std::string shape_key(const pgw_op* op) {
    return std::to_string(op->n) + "/" + std::to_string(op->k) + "/" + std::to_string(op->M) + "/" +
           std::to_string(op->Mp) + "/" + std::to_string(op->b) + "/" + std::to_string(op->nl) + "/" +
           (op->p == 0x7fffffffu ? "m31" : "bar");
}

// Tile configurations of a new job's operator: the cached choice for its
// shape, or the fastest of the groupable configurations 7-9 timed on this
// operator (its buffers are not yet in use by any job).
void tune_for_shape(pgw_batch* b, pgw_op* op) {
    const std::string key = shape_key(op);
    std::lock_guard<std::mutex> tl(b->tune_mtx);
    auto it = b->tuned.find(key);
    if (it == b->tuned.end()) {
        autotune_sweeps(op, 2, 7, 9, nullptr, nullptr);
        std::array<int, 4> c{};
        for (int w = 0; w < 4; ++w) c[w] = op->mfma_cfg[w];
        it = b->tuned.emplace(key, c).first;
    }
    for (int w = 0; w < 4; ++w) op->mfma_cfg[w] = it->second[w];
}

}  // namespace

extern "C" {

int pgw_batch_create(pgw_ctx* ctx, uint32_t flags, pgw_batch** out) {
    pgw_batch* b = nullptr;
    int rc = guarded([&] {
        require(ctx && out, "null context or output pointer");
        require((flags & ~(uint32_t)(PGW_BATCH_GRAPH | PGW_BATCH_GROUPED | PGW_BATCH_AUTOTUNE |
                                     PGW_BATCH_PERSISTENT)) == 0,
                "unknown batch flags");
        PGW_HIP(hipSetDevice(ctx->device));
        b = new pgw_batch;
        b->ctx = ctx;
        b->flags = flags;
        PGW_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        for (auto& u : b->upload) PGW_HIP(hipStreamCreateWithFlags(&u, hipStreamNonBlocking));
        ctx->live.fetch_add(1);
        *out = b;
    });
    if (rc != PGW_OK && b) {
        if (b->stream) (void)hipStreamDestroy(b->stream);
        for (auto u : b->upload) if (u) (void)hipStreamDestroy(u);
        delete b;
    }
    return rc;
}

int pgw_batch_add(pgw_batch* b, pgw_op* op, const uint32_t* h_v, uint64_t max_len, uint64_t* job) {
    BatchJob j;
    auto busy = [&] {
        for (const auto& kv : b->jobs)
            if (!kv.second.done && kv.second.op == op) return true;
        return false;
    };
    int rc = guarded([&] {
        require(b && op && h_v && job, "null batch, operator or buffer");
        require(op->ctx == b->ctx, "operator belongs to another context");
        require(max_len >= 1, "sequence length must be positive");
        {
            std::lock_guard<std::mutex> lk(b->mtx);
            require(!busy(), "an operator can carry one active job at a time");
        }
        const uint64_t N = op->cols, bb = op->b, ut = bb * (bb + 1) / 2;
        uint32_t vlo, vhi;
        min_max_u32(h_v, N * bb, vlo, vhi);
        require(N * bb == 0 || vhi < op->p, "V entries must be reduced modulo p");
        j.op = op;
        j.rounds_total = (max_len - 1 + 3) / 4;
        j.L = 1 + 4 * j.rounds_total;
        PGW_HIP(hipSetDevice(b->ctx->device));
        if ((b->flags & PGW_BATCH_AUTOTUNE) && op->kind == PGW_KIND_SYNTHETIC && op->method == PGW_METHOD_MFMA)
            tune_for_shape(b, op);
        {
            std::lock_guard<std::mutex> cl(b->ctx->capture_mtx);
            PGW_HIP(hipMalloc(&j.d_B, N * bb * sizeof(uint32_t)));
            PGW_HIP(hipMalloc(&j.d_D, N * bb * sizeof(uint32_t)));
            PGW_HIP(hipMalloc(&j.d_seq, j.L * ut * sizeof(uint32_t)));
            PGW_HIP(hipMalloc(&j.d_pos, sizeof(uint32_t)));
            static const uint32_t kFirstPosition = 1;  // seq[0] = V^T V below
            PGW_HIP(hipMemcpy(j.d_pos, &kFirstPosition, sizeof(uint32_t), hipMemcpyHostToDevice));
        }
        // staged on one of the batch's upload streams: steps of other jobs keep running
        const unsigned u = b->next_upload.fetch_add(1) % kUploadStreams;
        std::lock_guard<std::mutex> ul(b->upload_mtx[u]);
        std::lock_guard<std::mutex> cl(b->ctx->capture_mtx);
        hipStream_t up = b->upload[u];
        PGW_HIP(hipMemcpyAsync(j.d_B, h_v, N * bb * sizeof(uint32_t), hipMemcpyHostToDevice, up));
        try {
            launch_gram(j.d_B, j.d_B, N, op->b, 0, j.d_seq, nullptr, op->d_gram, op->m, up, nullptr, 0, op->gram_mfma);  // V^T V
        } catch (...) {
            (void)hipStreamSynchronize(up);  // the copy still reads h_v
            throw;
        }
        PGW_HIP(hipStreamSynchronize(up));
        if (j.rounds_total == 0) j.done = true;
        std::lock_guard<std::mutex> lk(b->mtx);
        require(!busy(), "an operator can carry one active job at a time");
        const uint64_t id = b->next_id++;
        j.id = id;
        b->jobs.emplace(id, j);
        *job = id;
    });
    if (rc != PGW_OK && (j.d_B || j.d_D || j.d_seq || j.d_pos)) {
        std::lock_guard<std::mutex> cl(b->ctx->capture_mtx);
        j.release();
    }
    return rc;
}

int pgw_batch_step(pgw_batch* b, uint32_t max_replays, uint64_t* active) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        PGW_HIP(hipSetDevice(b->ctx->device));
        const uint64_t a = b->step(max_replays);
        if (active) *active = a;
    });
}

int pgw_batch_poll(pgw_batch* b, uint64_t job, int* done) {
    return guarded([&] {
        require(b && done, "null batch or output");
        std::lock_guard<std::mutex> lk(b->mtx);
        auto it = b->jobs.find(job);
        require(it != b->jobs.end(), "unknown batch job");
        *done = it->second.done ? 1 : 0;
    });
}

int pgw_batch_result(pgw_batch* b, uint64_t job, uint32_t* h_seq, uint32_t* h_curv, uint64_t* len) {
    return guarded([&] {
        require(b && h_seq && h_curv, "null batch or buffer");
        // lock order: step_mtx, then the context's capture_mtx, then mtx
        std::lock_guard<std::mutex> sl(b->step_mtx);
        BatchJob j;
        {
            std::lock_guard<std::mutex> lk(b->mtx);
            auto it = b->jobs.find(job);
            require(it != b->jobs.end(), "unknown batch job");
            require(it->second.done, "batch job has not finished");
            j = it->second;
            b->jobs.erase(it);
        }
        // the job's buffers leave no live plan or graph behind them
        if (b->plan_uses(job)) b->drop_plan();
        std::lock_guard<std::mutex> cl(b->ctx->capture_mtx);
        struct Free {
            BatchJob& j;
            ~Free() { j.release(); }
        } guard{j};
        PGW_HIP(hipSetDevice(b->ctx->device));
        const uint64_t N = j.op->cols, bb = j.op->b, ut = bb * (bb + 1) / 2;
        PGW_HIP(hipMemcpy(h_seq, j.d_seq, j.L * ut * sizeof(uint32_t), hipMemcpyDeviceToHost));
        PGW_HIP(hipMemcpy(h_curv, j.d_B, N * bb * sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (len) *len = j.L;
    });
}

int pgw_batch_terms(pgw_batch* b, uint64_t job, uint64_t first, uint32_t* h_seq, uint64_t* done) {
    return guarded([&] {
        require(b && done, "null batch or output");
        uint32_t* d_seq = nullptr;
        uint64_t ut = 0, upto = 0;
        {
            std::lock_guard<std::mutex> lk(b->mtx);
            auto it = b->jobs.find(job);
            require(it != b->jobs.end(), "unknown batch job");
            const BatchJob& j = it->second;
            ut = (uint64_t)j.op->b * (j.op->b + 1) / 2;
            // rounds_done counts replays whose stream was synchronised
            upto = std::min<uint64_t>(j.L, 1 + 4 * j.rounds_done);
            d_seq = j.d_seq;
        }
        *done = upto;
        if (!h_seq || first >= upto) return;
        // on an upload stream, so that a step in progress is not waited for
        const unsigned u = b->next_upload.fetch_add(1) % kUploadStreams;
        std::lock_guard<std::mutex> ul(b->upload_mtx[u]);
        std::lock_guard<std::mutex> cl(b->ctx->capture_mtx);
        PGW_HIP(hipSetDevice(b->ctx->device));
        PGW_HIP(hipMemcpyAsync(h_seq + first * ut, d_seq + first * ut, (upto - first) * ut * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, b->upload[u]));
        PGW_HIP(hipStreamSynchronize(b->upload[u]));
    });
}

int pgw_batch_cancel(pgw_batch* b, uint64_t job) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        std::lock_guard<std::mutex> sl(b->step_mtx);
        PGW_HIP(hipSetDevice(b->ctx->device));
        PGW_HIP(hipStreamSynchronize(b->stream));
        {
            std::lock_guard<std::mutex> lk(b->mtx);
            require(b->jobs.find(job) != b->jobs.end(), "unknown batch job");
        }
        if (b->plan_uses(job)) b->drop_plan();
        std::lock_guard<std::mutex> cl(b->ctx->capture_mtx);
        std::lock_guard<std::mutex> lk(b->mtx);
        auto it = b->jobs.find(job);
        it->second.release();
        b->jobs.erase(it);
    });
}

int pgw_batch_graph_builds(pgw_batch* b, uint64_t* builds) {
    return guarded([&] {
        require(b && builds, "null batch or output");
        std::lock_guard<std::mutex> sl(b->step_mtx);
        *builds = b->graph_builds;
    });
}

int pgw_batch_groups(pgw_batch* b, uint64_t* groups, uint64_t* grouped_jobs) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        std::lock_guard<std::mutex> sl(b->step_mtx);
        uint64_t n = 0;
        for (const Group& g : b->groups) n += g.ids.size();
        if (groups) *groups = b->groups.size();
        if (grouped_jobs) *grouped_jobs = n;
    });
}

int pgw_batch_destroy(pgw_batch* b) {
    return guarded([&] {
        if (!b) return;
        {
            std::lock_guard<std::mutex> sl(b->step_mtx);
            PGW_HIP(hipSetDevice(b->ctx->device));
            (void)hipStreamSynchronize(b->stream);
            for (auto u : b->upload) (void)hipStreamSynchronize(u);
            b->drop_plan();
            {
                std::lock_guard<std::mutex> cl(b->ctx->capture_mtx);
                for (auto& kv : b->jobs) kv.second.release();
            }
            (void)hipStreamDestroy(b->stream);
            for (auto u : b->upload) (void)hipStreamDestroy(u);
        }
        b->ctx->live.fetch_sub(1);
        delete b;
    });
}

}  // extern "C"

// ===========================================================================
// serving: admission by HBM footprint, one continuously refilled batch
// ===========================================================================

namespace {

using Clock = std::chrono::steady_clock;

// Owned copy of an operator description and the arrays it points to.
struct OwnedDesc {
    pgw_operator_desc d{};
    std::vector<uint32_t> table_d, table_e, mu, right_inverse, kernel_basis, dense;
    std::vector<uint16_t> wv, wa0, wa1, kw;

    void copy_from(const pgw_operator_desc& s) {
        d = s;
        auto grab = [](auto& dst, const auto* src, uint64_t cnt) {
            if (src && cnt) dst.assign(src, src + cnt);
        };
        if (s.kind == PGW_KIND_SYNTHETIC) {
            const uint64_t cnt = (uint64_t)s.n * s.m_tgt * s.m_src;
            grab(table_d, s.table_d, cnt);
            grab(table_e, s.table_e, cnt);
            d.table_d = table_d.data();
            d.table_e = table_e.data();
        } else {
            const uint64_t n = s.genus - 3;
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
            if (s.kind == PGW_KIND_CYCLIC_AUGMENTED || s.kind == PGW_KIND_CYCLIC_REPLACEMENT) {
                uint64_t rows = 0;
                if (pgw_operator_dims(&s, &rows, nullptr) != PGW_OK) throw Error(PGW_ERR_INVALID, pgw_last_error());
                grab(dense, s.dense_block, rows * s.dense_cols);
                d.dense_block = dense.data();
            }
        }
    }
};

enum JobState { QUEUED = 0, RUNNING = 1, DONE = 2, FAILED = 3 };

constexpr int kServerWorkers = 4;  // host threads for set-up and result copies

struct ServerJob {
    uint64_t ticket = 0;
    OwnedDesc desc;
    std::vector<uint32_t> dr, dc, v;
    std::vector<pgw_rowmix_pair> pairs;
    uint32_t rounds = 0;
    uint64_t max_len = 0, rows = 0, cols = 0, charge = 0;
    pgw_op* op = nullptr;
    uint64_t batch_job = 0;
    int state = QUEUED;
    std::string error;
    std::vector<uint32_t> seq, curv;
    uint64_t L = 0;
    Clock::time_point t_submit, t_admit, t_started, t_left, t_finish;
    // host rank stage (pgw_server_enable_rank)
    bool rank = false, keep_generator = false, ranked = false;
    int rank_rc = PGW_OK;
    std::string rank_error;
    pgw_rank_info rank_info{};
    std::vector<uint32_t> gen;
    uint64_t gen_terms = 0;
    double rank_s = 0;
    // the rank stage fed while the job runs (no generator asked for): the
    // consumer's state, the terms it was given, and the feed in progress
    pgw_rank* rk = nullptr;
    uint64_t fed = 0;
    int feed = 0;  // 0 none, 1 queued, 2 running
};

// Terms a running job must have gained before the rank stage is fed again:
// at least this many, and at least as many as it was given so far, so the
// incremental steps stay few (about log2(L / kFeedMin)) and the consumer's
// total work stays within about twice one pass over the whole sequence.
constexpr uint64_t kFeedMin = 512;

// Device bytes of a job: operator (tables, digit planes, intermediate blocks,
// preconditioners, row-mix plan, Gram scratch) plus its sequence state.
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
        bytes += 2 * us * 4 * b * (((uint64_t)d.m_src + 31) / 32 * 32);   // source digit planes
        bytes += 2 * ut2 * 4 * b * (((uint64_t)d.m_tgt + 31) / 32 * 32);  // mid digit planes
        bytes += (us + ut2) * ((uint64_t)d.n + 1) * 4;                    // incidence tables
        bytes += cols * b * 4 + (cols + (cols / 2 + 1) * 2 * kMixHalo) * 17;  // row-mixing chains
    } else {
        // y', the mixed target and the five sparse maps of the chain (sliced
        // ELL, 8 bytes per entry; rows of the paper instances hold < 100
        // entries per map, padding included), the row-mixing order of t
        bytes += 3 * (rows + cols + 2) * b * 4 + 5 * 100 * 8 * (rows + cols + 64) + 8 * rows;
    }
    return bytes + (16ull << 20);
}

}  // namespace

struct pgw_server {
    pgw_ctx* ctx = nullptr;
    uint64_t budget = 0;
    uint32_t max_active = 0, replays_per_step = 1;
    pgw_batch* batch = nullptr;
    std::mutex mtx;
    std::condition_variable cv_io, cv_step, cv_done, cv_feed;
    std::deque<uint64_t> queue;                              // submitted, not yet admitted
    std::deque<uint64_t> feeds;                              // running jobs with terms for the rank stage
    std::vector<uint64_t> running;                           // in the batch
    std::deque<std::pair<uint64_t, std::string>> finished;   // left the batch (error text if failed)
    std::map<uint64_t, std::unique_ptr<ServerJob>> jobs;
    uint64_t next_ticket = 1, in_use = 0, peak_bytes = 0, peak_active = 0, completed = 0;
    uint32_t starting = 0;  // admitted jobs still being set up
    bool stop = false, step_exit = false;
    bool rank_on = false, rank_keep_gen = false;
    uint32_t rank_threads = 1, rank_threshold = PGW_RANK_DEFAULT_THRESHOLD,
             rank_buffer = PGW_RANK_DEFAULT_MAX_NLEN_BUFFER;
    std::vector<std::thread> workers;
    std::thread stepper;
    Clock::time_point t0 = Clock::now();

    // under mtx
    bool can_admit() const {
        if (queue.empty() || running.size() + starting >= max_active) return false;
        return in_use + jobs.at(queue.front())->charge <= budget;
    }

    // under mtx
    void finish(ServerJob& j, int state, const std::string& err) {
        j.state = state;
        j.error = err;
        j.t_finish = Clock::now();
        ++completed;
        cv_done.notify_all();
    }

    // Host workers: each takes one finished job to collect or one queued job
    // to start (operator set-up, uploads), so host work runs beside the GPU
    // steps and several jobs' host work runs at once.
    void io_loop() {
        (void)hipSetDevice(ctx->device);
        for (;;) {
            ServerJob* admit = nullptr;
            ServerJob* feed = nullptr;
            std::pair<uint64_t, std::string> fin{0, ""};
            {
                std::unique_lock<std::mutex> lk(mtx);
                cv_io.wait(lk, [&] {
                    return !finished.empty() || can_admit() || !feeds.empty() ||
                           (stop && queue.empty() && running.empty() && starting == 0);
                });
                if (!finished.empty()) {
                    fin = finished.front();
                    finished.pop_front();
                } else if (!feeds.empty() && !can_admit()) {
                    feed = jobs.at(feeds.front()).get();
                    feeds.pop_front();
                    feed->feed = 2;
                } else if (can_admit()) {
                    admit = jobs.at(queue.front()).get();
                    queue.pop_front();
                    in_use += admit->charge;
                    peak_bytes = std::max(peak_bytes, in_use);
                    admit->state = RUNNING;
                    ++starting;
                } else {
                    cv_io.notify_all();  // stopping with nothing left: wake the other workers
                    return;
                }
            }
            if (fin.first) collect(fin.first, fin.second);
            if (admit) start(*admit);
            if (feed) feed_rank(*feed);
        }
    }

    void start(ServerJob& j) {
        j.t_admit = Clock::now();
        int rc = pgw_operator_create(ctx, &j.desc.d, j.dr.data(), j.dc.data(), j.rounds,
                                     j.pairs.empty() ? nullptr : j.pairs.data(), &j.op);
        if (rc == PGW_OK) rc = pgw_batch_add(batch, j.op, j.v.data(), j.max_len, &j.batch_job);
        std::string err = rc == PGW_OK ? "" : pgw_last_error();
        uint64_t actual = 0;
        if (rc == PGW_OK) {
            // charge what the operator and the job really hold
            pgw_op_info info{};
            pgw_operator_info(j.op, &info);
            const uint64_t b = j.desc.d.block, ut = b * (b + 1) / 2;
            const uint64_t L = 1 + 4 * ((j.max_len - 1 + 3) / 4);
            actual = info.device_bytes + 2 * j.cols * b * 4 + L * ut * 4;
        } else if (j.op) {
            pgw_operator_destroy(j.op);
            j.op = nullptr;
        }
        j.dr = std::vector<uint32_t>();
        j.dc = std::vector<uint32_t>();
        j.v = std::vector<uint32_t>();
        j.pairs = std::vector<pgw_rowmix_pair>();
        std::lock_guard<std::mutex> lk(mtx);
        --starting;
        if (rc != PGW_OK) {
            in_use -= j.charge;
            j.t_started = j.t_left = Clock::now();
            finish(j, FAILED, err);
            cv_io.notify_all();
            return;
        }
        in_use = in_use - j.charge + actual;
        j.charge = actual;
        j.t_started = Clock::now();
        if (j.rank && !j.keep_generator) {
            // fed during the job; a failure here leaves the one-pass stage at the end
            const uint64_t b = j.desc.d.block, ut = b * (b + 1) / 2;
            j.seq.resize((1 + 4 * ((j.max_len - 1 + 3) / 4)) * ut);
            if (pgw_rank_create(j.desc.d.prime, j.desc.d.block, j.rows, j.cols, rank_threshold, rank_buffer,
                                rank_threads, &j.rk) != PGW_OK)
                j.rk = nullptr;
        }
        peak_bytes = std::max(peak_bytes, in_use);
        running.push_back(j.ticket);
        peak_active = std::max<uint64_t>(peak_active, running.size());
        cv_step.notify_all();
    }

    void collect(uint64_t ticket, const std::string& err) {
        ServerJob* j;
        {
            // a queued feed is dropped, a running one finished first: the
            // result below releases the job's device sequence
            std::unique_lock<std::mutex> lk(mtx);
            j = jobs.at(ticket).get();
            if (j->feed == 1) {
                feeds.erase(std::find(feeds.begin(), feeds.end(), ticket));
                j->feed = 0;
            }
            cv_feed.wait(lk, [&] { return j->feed == 0; });
        }
        int rc = PGW_OK;
        std::string msg = err;
        if (err.empty()) {
            const uint64_t b = j->desc.d.block, ut = b * (b + 1) / 2;
            j->seq.resize((1 + 4 * ((j->max_len - 1 + 3) / 4)) * ut);
            j->curv.resize(j->cols * b);
            rc = pgw_batch_result(batch, j->batch_job, j->seq.data(), j->curv.data(), &j->L);
            if (rc != PGW_OK) msg = pgw_last_error();
        } else {
            pgw_batch_cancel(batch, j->batch_job);
            rc = PGW_ERR_HIP;
        }
        pgw_operator_destroy(j->op);
        j->op = nullptr;
        {
            // the device memory is free: admissions need not wait for the rank stage
            std::lock_guard<std::mutex> lk(mtx);
            in_use -= j->charge;
            cv_io.notify_all();
        }
        if (rc == PGW_OK && j->rank) rank_job(*j);
        if (j->rk) {
            pgw_rank_destroy(j->rk);
            j->rk = nullptr;
        }
        std::lock_guard<std::mutex> lk(mtx);
        finish(*j, rc == PGW_OK ? DONE : FAILED, msg);
        cv_io.notify_all();
    }

    // RankComputation over the job's whole sequence (third_party/bcw/src/main.rs:
    // 46-57 on a completed file), on this worker thread
    void rank_job(ServerJob& j) {
        const Clock::time_point t = Clock::now();
        if (j.rk) {
            // the terms after those fed while the job ran
            int rc = j.fed < j.L ? pgw_rank_step(j.rk, j.seq.data(), j.L, &j.rank_info)
                                 : pgw_rank_get_info(j.rk, &j.rank_info);
            if (rc != PGW_OK) j.rank_error = pgw_last_error();
            j.rank_rc = rc;
            j.ranked = true;
            j.rank_s += std::chrono::duration<double>(Clock::now() - t).count();
            return;
        }
        pgw_rank* r = nullptr;
        int rc = pgw_rank_create(j.desc.d.prime, j.desc.d.block, j.rows, j.cols, rank_threshold,
                                 rank_buffer, rank_threads, &r);
        if (rc == PGW_OK) rc = pgw_rank_step(r, j.seq.data(), j.L, &j.rank_info);
        if (rc == PGW_OK && j.keep_generator) {
            rc = pgw_rank_generator(r, nullptr, 0, &j.gen_terms);
            if (rc == PGW_OK) {
                j.gen.resize(j.gen_terms * j.desc.d.block * j.desc.d.block);
                rc = pgw_rank_generator(r, j.gen.data(), j.gen.size(), &j.gen_terms);
            }
        }
        if (rc != PGW_OK) j.rank_error = pgw_last_error();
        j.rank_rc = rc;
        pgw_rank_destroy(r);
        j.ranked = true;
        j.rank_s = std::chrono::duration<double>(Clock::now() - t).count();
    }

    // One incremental rank step on the terms the job finished since the last
    // feed (RankComputation::step, third_party/bcw/src/rank_pipeline.rs:
    // 153-222, which the reference runs once on a completed file).  A
    // failure drops the streamed state: the job's rank stage then runs in
    // one pass after it leaves the GPU.
    void feed_rank(ServerJob& j) {
        const Clock::time_point t = Clock::now();
        uint64_t done = 0;
        int rc = pgw_batch_terms(batch, j.batch_job, j.fed, j.seq.data(), &done);
        if (rc == PGW_OK && done > j.fed) {
            pgw_rank_info info{};
            rc = pgw_rank_step(j.rk, j.seq.data(), done, &info);
        }
        const double secs = std::chrono::duration<double>(Clock::now() - t).count();
        std::lock_guard<std::mutex> lk(mtx);
        j.rank_s += secs;
        if (rc == PGW_OK) {
            j.fed = std::max(j.fed, done);
        } else {
            pgw_rank_destroy(j.rk);
            j.rk = nullptr;
        }
        j.feed = 0;
        cv_feed.notify_all();
    }

    // the GPU thread: batch steps back to back while jobs are running
    void step_loop() {
        (void)hipSetDevice(ctx->device);
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mtx);
                cv_step.wait(lk, [&] { return !running.empty() || step_exit; });
                if (running.empty()) return;
            }
            const int rc = pgw_batch_step(batch, replays_per_step, nullptr);
            const std::string err = rc == PGW_OK ? "" : std::string(pgw_last_error());
            std::lock_guard<std::mutex> lk(mtx);
            std::vector<uint64_t> still;
            for (uint64_t t : running) {
                int done = 0;
                if (rc == PGW_OK) pgw_batch_poll(batch, jobs.at(t)->batch_job, &done);
                if (rc != PGW_OK || done) {
                    jobs.at(t)->t_left = Clock::now();
                    finished.emplace_back(t, err);
                }
                else
                    still.push_back(t);
            }
            running.swap(still);
            // running jobs whose finished terms have grown enough feed the rank stage
            for (uint64_t t : running) {
                ServerJob& j = *jobs.at(t);
                uint64_t done = 0;
                if (!j.rk || j.feed != 0 || pgw_batch_terms(batch, j.batch_job, 0, nullptr, &done) != PGW_OK)
                    continue;
                if (done > j.fed && done - j.fed >= std::max(kFeedMin, j.fed)) {
                    j.feed = 1;
                    feeds.push_back(t);
                }
            }
            cv_io.notify_all();
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
                      uint32_t replays_per_step, pgw_server** out) {
    pgw_server* s = nullptr;
    int rc = guarded([&] {
        require(ctx && out, "null context or output pointer");
        require(hbm_budget > 0 && max_active > 0, "budget and slot count must be positive");
        s = new pgw_server;
        s->ctx = ctx;
        s->budget = hbm_budget;
        s->max_active = max_active;
        s->replays_per_step = replays_per_step ? replays_per_step : 1;
        if (pgw_batch_create(ctx, PGW_BATCH_GRAPH | PGW_BATCH_GROUPED, &s->batch) != PGW_OK)
            throw Error(PGW_ERR_HIP, pgw_last_error());
        for (int w = 0; w < kServerWorkers; ++w) s->workers.emplace_back([s] { s->io_loop(); });
        s->stepper = std::thread([s] { s->step_loop(); });
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
        j->charge = estimate_job_bytes(*desc, rows, cols, max_len, rowmix_rounds);
        std::lock_guard<std::mutex> lk(s->mtx);
        require(!s->stop, "server is shutting down");
        j->rank = s->rank_on;
        j->keep_generator = s->rank_keep_gen;
        j->ticket = s->next_ticket++;
        *ticket = j->ticket;
        ServerJob& ref = *j;
        s->jobs.emplace(j->ticket, std::move(j));
        if (ref.charge > s->budget) {
            // can never fit: fail at once instead of blocking the queue
            ref.t_admit = ref.t_started = ref.t_left = ref.t_submit;
            s->finish(ref, FAILED, "job needs more device memory than the server budget");
            return;
        }
        s->queue.push_back(ref.ticket);
        s->cv_io.notify_all();
    });
}

int pgw_server_wait(pgw_server* s, uint64_t ticket, double timeout_s, int* state) {
    return guarded([&] {
        require(s && state, "null server or output");
        std::unique_lock<std::mutex> lk(s->mtx);
        auto it = s->jobs.find(ticket);
        require(it != s->jobs.end(), "unknown ticket");
        ServerJob& j = *it->second;
        auto pred = [&] { return j.state >= DONE; };
        if (timeout_s < 0) s->cv_done.wait(lk, pred);
        else s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
        *state = j.state;
    });
}

int pgw_server_result(pgw_server* s, uint64_t ticket, uint32_t* h_seq, uint32_t* h_curv,
                      uint64_t* len, double* times) {
    return guarded([&] {
        require(s, "null server");
        std::unique_lock<std::mutex> lk(s->mtx);
        auto it = s->jobs.find(ticket);
        require(it != s->jobs.end(), "unknown ticket");
        ServerJob& j = *it->second;
        require(j.state >= DONE, "job has not finished");
        if (j.state == FAILED) {
            const std::string msg = "job failed: " + j.error;
            s->jobs.erase(it);
            throw Error(PGW_ERR_INVALID, msg);
        }
        if (h_seq) std::copy(j.seq.begin(), j.seq.end(), h_seq);
        if (h_curv) std::copy(j.curv.begin(), j.curv.end(), h_curv);
        if (len) *len = j.L;
        if (times) {
            const Clock::time_point tp[5] = {j.t_submit, j.t_admit, j.t_started, j.t_left, j.t_finish};
            for (int q = 0; q < 5; ++q) times[q] = std::chrono::duration<double>(tp[q] - s->t0).count();
        }
        s->jobs.erase(it);
    });
}

int pgw_server_enable_rank(pgw_server* s, uint32_t threads, uint32_t threshold,
                           uint32_t max_nlen_buffer, int keep_generator) {
    return guarded([&] {
        require(s != nullptr, "null server");
        std::lock_guard<std::mutex> lk(s->mtx);
        s->rank_on = true;
        s->rank_threads = threads ? threads : 1;
        s->rank_threshold = threshold;
        s->rank_buffer = max_nlen_buffer;
        s->rank_keep_gen = keep_generator != 0;
    });
}

int pgw_server_rank(pgw_server* s, uint64_t ticket, pgw_rank_info* info, uint32_t* h_gen,
                    uint64_t gen_capacity, uint64_t* gen_terms, double* seconds) {
    return guarded([&] {
        require(s != nullptr, "null server");
        std::lock_guard<std::mutex> lk(s->mtx);
        auto it = s->jobs.find(ticket);
        require(it != s->jobs.end(), "unknown ticket");
        ServerJob& j = *it->second;
        require(j.state >= DONE, "job has not finished");
        require(j.state == DONE, "job failed: " + j.error);
        require(j.ranked, "the job was submitted without the rank stage");
        if (j.rank_rc != PGW_OK) throw Error(j.rank_rc, "rank stage: " + j.rank_error);
        if (info) *info = j.rank_info;
        if (seconds) *seconds = j.rank_s;
        if (gen_terms) *gen_terms = j.gen_terms;
        if (h_gen) {
            require(j.keep_generator, "the generator was not kept (keep_generator = 0)");
            require(gen_capacity >= j.gen.size(), "generator buffer too small: " +
                                                      std::to_string(j.gen.size()) + " values needed");
            std::copy(j.gen.begin(), j.gen.end(), h_gen);
        }
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
        s->cv_io.notify_all();
        for (auto& w : s->workers) w.join();
        {
            std::lock_guard<std::mutex> lk(s->mtx);
            s->step_exit = true;
        }
        s->cv_step.notify_all();
        if (s->stepper.joinable()) s->stepper.join();
        for (auto& kv : s->jobs) {
            if (kv.second->op) pgw_operator_destroy(kv.second->op);
            if (kv.second->rk) pgw_rank_destroy(kv.second->rk);
        }
        pgw_batch_destroy(s->batch);
        delete s;
    });
}

}  // extern "C"

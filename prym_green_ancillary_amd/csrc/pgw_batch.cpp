// pgw_batch.cpp -- batches of sector sequence loops advanced by replayed HIP
// graphs, and the serving loop that feeds them.
//
// The reference runs one sector per process, one after another
// (scripts/reproduce.py:235-276 around cuda/cuprym_cyclic.cu:2316-2461).  Here
// one process keeps many sectors resident in HBM and advances them together:
// small sectors cannot fill 256 CUs on their own, several of them can.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
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

namespace {

constexpr uint32_t kMaxQuota = 8;       // rounds of one job inside one graph replay
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
    double cost = 0.0;  // modelled seconds per round

    void release() {
        for (uint32_t* p : {d_B, d_D, d_seq, d_pos})
            if (p) (void)hipFree(p);
        d_B = d_D = d_seq = d_pos = nullptr;
    }
};

// Modelled device time of one round (two block-applies, four projections):
// table-contraction work at the rate each path reaches on a large sector,
// the elementwise passes at HBM speed, and a floor for ~24 kernel launches.
double round_cost(const pgw_op* op) {
    const double b = op->b;
    double s = 0.0;
    if (op->kind == PGW_KIND_SYNTHETIC) {
        const double macs = 4.0 * (double)op->u_tgt * (op->n - op->k + 1) * op->Mp * op->M * b;
        if (op->method == PGW_METHOD_MFMA)
            s = macs * op->nl * op->nl / 4.0e14;
        else
            s = macs / 4.0e12;
        s += 40.0 * (op->rows + op->cols + op->mid) * b * 4.0 / 4.0e12;
    } else {
        s = 200.0 * (op->rows + op->cols + op->mid) * b * 4.0 / 2.0e12;
    }
    return 2.0 * s + 100e-6;
}

// One round of the sequence loop for a job (cuprym_cyclic.cu:2439-2449):
// D = S B; seq[pos] = B^T D, seq[pos+1] = D^T D; B = S D; seq[pos+2] = B^T D,
// seq[pos+3] = B^T B; pos += 4.  Positions are read on the device, so the
// same launches (or graph) serve every round.
void enqueue_round(const BatchJob& j, hipStream_t st) {
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

struct Slot {
    uint64_t id;
    uint32_t quota;
    bool operator==(const Slot& o) const { return id == o.id && quota == o.quota; }
};

}  // namespace

struct pgw_batch {
    pgw_ctx* ctx = nullptr;
    uint32_t flags = 0;
    std::vector<hipStream_t> streams;
    hipStream_t upload[kUploadStreams] = {};
    std::mutex upload_mtx[kUploadStreams];
    std::atomic<unsigned> next_upload{0};
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;
    // `mtx` guards the job table; `step_mtx` serialises steps with cancel/destroy
    std::mutex mtx, step_mtx;
    std::map<uint64_t, BatchJob> jobs;
    uint64_t next_id = 1;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<Slot> graph_key;  // (job, rounds per replay) the graph was captured for
    uint64_t graph_builds = 0;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_key.clear();
    }

    // every slot's rounds, fanned out over the streams and joined on streams[0]
    void enqueue_all(const std::vector<Slot>& key, const std::vector<const BatchJob*>& js) {
        hipStream_t s0 = streams[0];
        const size_t S = std::min(streams.size(), js.size());
        PGW_HIP(hipEventRecord(fork, s0));
        for (size_t q = 1; q < S; ++q) PGW_HIP(hipStreamWaitEvent(streams[q], fork, 0));
        // largest jobs first, so that the longest chains start earliest
        for (size_t q = 0; q < js.size(); ++q)
            for (uint32_t r = 0; r < key[q].quota; ++r) enqueue_round(*js[q], streams[q % S]);
        for (size_t q = 1; q < S; ++q) {
            PGW_HIP(hipEventRecord(joins[q], streams[q]));
            PGW_HIP(hipStreamWaitEvent(s0, joins[q], 0));
        }
    }

    uint64_t step(uint32_t max_replays) {
        std::lock_guard<std::mutex> sl(step_mtx);
        std::vector<Slot> key;
        std::vector<BatchJob*> js;
        {
            std::lock_guard<std::mutex> lk(mtx);
            for (auto& kv : jobs)
                if (!kv.second.done) js.push_back(&kv.second);
            if (js.empty()) return 0;
            std::stable_sort(js.begin(), js.end(),
                             [](const BatchJob* a, const BatchJob* b) { return a->cost > b->cost; });
            const double top = js.front()->cost;
            for (BatchJob* j : js) {
                const uint64_t left = j->rounds_total - j->rounds_done;
                uint64_t q = 1;
                if (flags & PGW_BATCH_BALANCE) q = (uint64_t)std::floor(top / j->cost);
                q = std::max<uint64_t>(1, std::min<uint64_t>({q, kMaxQuota, left}));
                key.push_back({j->id, (uint32_t)q});
            }
        }
        // replays while every job still has a full quota left
        uint64_t reps = std::max<uint32_t>(1, max_replays);
        for (size_t q = 0; q < js.size(); ++q)
            reps = std::min<uint64_t>(reps, (js[q]->rounds_total - js[q]->rounds_done) / key[q].quota);
        std::vector<const BatchJob*> cjs(js.begin(), js.end());
        hipStream_t s0 = streams[0];
        if (flags & PGW_BATCH_GRAPH) {
            if (!(key == graph_key) || !exec) {
                std::lock_guard<std::mutex> cl(ctx->capture_mtx);
                drop_graph();
                PGW_HIP(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
                try {
                    enqueue_all(key, cjs);
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(s0, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                PGW_HIP(hipStreamEndCapture(s0, &graph));
                PGW_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
                graph_key = key;
                ++graph_builds;
            }
            for (uint64_t r = 0; r < reps; ++r) PGW_HIP(hipGraphLaunch(exec, s0));
        } else {
            for (uint64_t r = 0; r < reps; ++r) enqueue_all(key, cjs);
        }
        PGW_HIP(hipStreamSynchronize(s0));
        std::lock_guard<std::mutex> lk(mtx);
        uint64_t active = 0;
        for (size_t q = 0; q < js.size(); ++q) {
            js[q]->rounds_done += reps * key[q].quota;
            if (js[q]->rounds_done >= js[q]->rounds_total) js[q]->done = true;
        }
        for (auto& kv : jobs) active += kv.second.done ? 0 : 1;
        return active;
    }
};

extern "C" {

int pgw_batch_create(pgw_ctx* ctx, uint32_t nstreams, uint32_t flags, pgw_batch** out) {
    pgw_batch* b = nullptr;
    int rc = guarded([&] {
        require(ctx && out, "null context or output pointer");
        require((flags & ~(uint32_t)(PGW_BATCH_GRAPH | PGW_BATCH_BALANCE)) == 0, "unknown batch flags");
        PGW_HIP(hipSetDevice(ctx->device));
        b = new pgw_batch;
        b->ctx = ctx;
        b->flags = flags;
        if (nstreams == 0) nstreams = 8;
        b->streams.resize(nstreams, nullptr);
        b->joins.resize(nstreams, nullptr);
        for (uint32_t q = 0; q < nstreams; ++q) {
            PGW_HIP(hipStreamCreateWithFlags(&b->streams[q], hipStreamNonBlocking));
            PGW_HIP(hipEventCreateWithFlags(&b->joins[q], hipEventDisableTiming));
        }
        for (auto& u : b->upload) PGW_HIP(hipStreamCreateWithFlags(&u, hipStreamNonBlocking));
        PGW_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
        ctx->live.fetch_add(1);
        *out = b;
    });
    if (rc != PGW_OK && b) {
        for (auto s : b->streams) if (s) (void)hipStreamDestroy(s);
        for (auto e : b->joins) if (e) (void)hipEventDestroy(e);
        for (auto u : b->upload) if (u) (void)hipStreamDestroy(u);
        if (b->fork) (void)hipEventDestroy(b->fork);
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
        j.cost = round_cost(op);
        PGW_HIP(hipSetDevice(b->ctx->device));
        {
            std::lock_guard<std::mutex> cl(b->ctx->capture_mtx);
            PGW_HIP(hipMalloc(&j.d_B, N * bb * sizeof(uint32_t)));
            PGW_HIP(hipMalloc(&j.d_D, N * bb * sizeof(uint32_t)));
            PGW_HIP(hipMalloc(&j.d_seq, j.L * ut * sizeof(uint32_t)));
            PGW_HIP(hipMalloc(&j.d_pos, sizeof(uint32_t)));
        }
        // staged on one of the batch's upload streams: steps of other jobs keep running
        const unsigned u = b->next_upload.fetch_add(1) % kUploadStreams;
        std::lock_guard<std::mutex> ul(b->upload_mtx[u]);
        std::lock_guard<std::mutex> cl(b->ctx->capture_mtx);
        hipStream_t up = b->upload[u];
        const uint32_t one = 1;
        PGW_HIP(hipMemcpyAsync(j.d_B, h_v, N * bb * sizeof(uint32_t), hipMemcpyHostToDevice, up));
        PGW_HIP(hipMemcpyAsync(j.d_pos, &one, sizeof(uint32_t), hipMemcpyHostToDevice, up));
        launch_gram(j.d_B, j.d_B, N, op->b, 0, j.d_seq, nullptr, op->d_gram, op->m, up);  // V^T V
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
        BatchJob j;
        {
            std::lock_guard<std::mutex> lk(b->mtx);
            auto it = b->jobs.find(job);
            require(it != b->jobs.end(), "unknown batch job");
            require(it->second.done, "batch job has not finished");
            j = it->second;
            b->jobs.erase(it);
        }
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

int pgw_batch_cancel(pgw_batch* b, uint64_t job) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        // lock order: step_mtx, then the context's capture_mtx, then mtx
        std::lock_guard<std::mutex> sl(b->step_mtx);
        std::lock_guard<std::mutex> cl(b->ctx->capture_mtx);
        std::lock_guard<std::mutex> lk(b->mtx);
        auto it = b->jobs.find(job);
        require(it != b->jobs.end(), "unknown batch job");
        PGW_HIP(hipSetDevice(b->ctx->device));
        for (auto s : b->streams) PGW_HIP(hipStreamSynchronize(s));
        b->drop_graph();
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

int pgw_batch_destroy(pgw_batch* b) {
    return guarded([&] {
        if (!b) return;
        {
            std::lock_guard<std::mutex> sl(b->step_mtx);
            PGW_HIP(hipSetDevice(b->ctx->device));
            for (auto s : b->streams) (void)hipStreamSynchronize(s);
            for (auto u : b->upload) (void)hipStreamSynchronize(u);
            b->drop_graph();
            for (auto& kv : b->jobs) kv.second.release();
            for (auto s : b->streams) (void)hipStreamDestroy(s);
            for (auto e : b->joins) (void)hipEventDestroy(e);
            for (auto u : b->upload) (void)hipStreamDestroy(u);
            (void)hipEventDestroy(b->fork);
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
};

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
    } else {
        // y', the mixed target and the five sparse maps of the chain (sliced
        // ELL, 8 bytes per entry; rows of the paper instances hold < 100
        // entries per map, padding included)
        bytes += 3 * (rows + cols + 2) * b * 4 + 5 * 100 * 8 * (rows + cols + 64);
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
    std::condition_variable cv_io, cv_step, cv_done;
    std::deque<uint64_t> queue;                              // submitted, not yet admitted
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
            std::pair<uint64_t, std::string> fin{0, ""};
            {
                std::unique_lock<std::mutex> lk(mtx);
                cv_io.wait(lk, [&] {
                    return !finished.empty() || can_admit() ||
                           (stop && queue.empty() && running.empty() && starting == 0);
                });
                if (!finished.empty()) {
                    fin = finished.front();
                    finished.pop_front();
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
        peak_bytes = std::max(peak_bytes, in_use);
        running.push_back(j.ticket);
        peak_active = std::max<uint64_t>(peak_active, running.size());
        cv_step.notify_all();
    }

    void collect(uint64_t ticket, const std::string& err) {
        ServerJob* j;
        {
            std::lock_guard<std::mutex> lk(mtx);
            j = jobs.at(ticket).get();
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
        std::lock_guard<std::mutex> lk(mtx);
        finish(*j, rc == PGW_OK ? DONE : FAILED, msg);
        cv_io.notify_all();
    }

    // RankComputation over the job's whole sequence (third_party/bcw/src/main.rs:
    // 46-57 on a completed file), on this worker thread
    void rank_job(ServerJob& j) {
        const Clock::time_point t = Clock::now();
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
        if (pgw_batch_create(ctx, std::min<uint32_t>(max_active, 8),
                             PGW_BATCH_GRAPH | PGW_BATCH_BALANCE, &s->batch) != PGW_OK)
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
        for (auto& kv : s->jobs)
            if (kv.second->op) pgw_operator_destroy(kv.second->op);
        pgw_batch_destroy(s->batch);
        delete s;
    });
}

}  // extern "C"

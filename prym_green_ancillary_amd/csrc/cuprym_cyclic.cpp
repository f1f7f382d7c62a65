// cuprym_cyclic -- sector sequence generator (the reference's cuprym_cyclic
// executable, cuda/cuprym_cyclic.cu:2498-2689), driving libpgw through its C ABI.
//
//   cuprym_cyclic INSTANCE --sector S [--operator eliminated|augmented|replacement|synthetic]
//       [--dense-file F] [--drop-columns F] [--out DIR]
//       [-d DEV] [-v B] [-s 0] [--seed 1] [--rowmix-rounds 2] [--rowmix-seed 10001]
//       [--validate-rowmix] [--length L] [--extra-terms 50] [--method auto|plain|mfma]
//       [--no-save]
//
// Same option names, output names (DIR/<op>-sectorSS-vB-rowmix-rR-sS.wdm.zst
// plus .rowmix.json), WDM text and exit codes as the reference; the values the
// reference pins to the paper recipe (v = 4, p < 2^16) are relaxed: 1 <= b <= 16,
// p < 2^31, and a synthetic sector spec is accepted besides the cyclic instance.

#include <chrono>
#include <cstdio>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <memory>
#include <vector>

#include "cli_common.h"
#include "wdm.h"

using namespace pgwcli;

namespace {

uint64_t dot_mod(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b, uint32_t p) {
    unsigned __int128 s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += (uint64_t)a[i] * b[i];
    return (uint64_t)(s % p);
}

// host restatement of one row-mixing pass, used only by --validate-rowmix
void host_rowmix(std::vector<uint32_t>& x, const std::vector<pgw_rowmix_pair>& pairs,
                 uint64_t pc, uint32_t rounds, uint32_t b, uint32_t p, bool transpose) {
    for (uint32_t t = 0; t < rounds; ++t) {
        const uint32_t r = transpose ? rounds - 1 - t : t;
        for (uint64_t q = 0; q < pc; ++q) {
            const pgw_rowmix_pair& pr = pairs[(uint64_t)r * pc + q];
            const uint64_t opab = (1 + (uint64_t)pr.a * pr.b % p) % p;
            for (uint32_t c = 0; c < b; ++c) {
                uint64_t xv = x[(uint64_t)pr.first * b + c], yv = x[(uint64_t)pr.second * b + c];
                uint64_t f = transpose ? (xv + (uint64_t)pr.b * yv) % p : (xv + (uint64_t)pr.a * yv) % p;
                uint64_t s = transpose ? ((uint64_t)pr.a * xv % p + opab * yv % p) % p
                                       : ((uint64_t)pr.b * xv % p + opab * yv % p) % p;
                x[(uint64_t)pr.first * b + c] = (uint32_t)f;
                x[(uint64_t)pr.second * b + c] = (uint32_t)s;
            }
        }
    }
}

// the loop's progress: a status line, and the finished terms handed to the
// streaming WDM writer (rows [0, done) of the sequence are final here)
struct LoopState {
    std::chrono::steady_clock::time_point start;
    pgwio::WdmStream* stream = nullptr;
    const uint32_t* seq = nullptr;
};

void progress_cb(uint64_t done, uint64_t total, void* user) {
    auto* st = static_cast<LoopState*>(user);
    if (st->stream) st->stream->terms(st->seq, done);
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - st->start).count();
    double rate = secs > 0 ? done / secs : 0.0;
    std::cout << "\rProgress: " << done << "/" << total << " | Elapsed: " << (long long)secs
              << "s | Throughput: " << std::fixed << std::setprecision(2) << rate << "/s"
              << " | Remaining: " << std::setprecision(1) << (rate > 0 ? (total - done) / rate : 0.0)
              << "s      " << std::flush;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const std::vector<OptSpec> specs = {
            {{"--operator"}, false}, {{"--sector"}, false}, {{"--out"}, false},
            {{"-d", "--device"}, false}, {{"-v"}, false}, {{"-s", "--saveafter"}, false},
            {{"--seed"}, false}, {{"--rowmix-rounds"}, false}, {{"--rowmix-seed"}, false},
            {{"--validate-rowmix"}, true}, {{"--length"}, false}, {{"--extra-terms"}, false},
            {{"--method"}, false}, {{"--no-save"}, true}, {{"-h", "--help"}, true},
            {{"--dense-file"}, false}, {{"--drop-columns"}, false},
        };
        Args a = parse_args(argc, argv, specs);
        if (has(a, "--help") || a.positional.size() != 1 || !has(a, "--sector")) {
            std::cerr << "usage: cuprym_cyclic INSTANCE --sector S "
                         "[--operator eliminated|augmented|replacement|synthetic] [--dense-file F] "
                         "[--drop-columns F] [--out DIR] [-d DEV] [-v B] [--seed N] [--rowmix-rounds R] "
                         "[--rowmix-seed N] [--validate-rowmix] [--length L] [--extra-terms E] "
                         "[--method auto|plain|mfma] [--no-save]\n";
            return has(a, "--help") ? 0 : 1;
        }
        const std::string instance_path = a.positional[0];
        const std::string out_dir = get(a, "--out", ".");
        const uint32_t sector = (uint32_t)get_u64(a, "--sector", 0);
        const int device = (int)get_u64(a, "--device", 0);
        const uint32_t num_v = (uint32_t)get_u64(a, "-v", 4);
        const uint64_t save_after = get_u64(a, "--saveafter", 0);
        const uint64_t seed = get_u64(a, "--seed", 1);
        const uint32_t rounds = (uint32_t)get_u64(a, "--rowmix-rounds", 2);
        const uint64_t rowmix_seed = get_u64(a, "--rowmix-seed", 10001);
        const bool no_save = has(a, "--no-save");
        const std::string method_name = get(a, "--method", "auto");
        uint32_t method = PGW_METHOD_AUTO;
        if (method_name == "plain") method = PGW_METHOD_PLAIN;
        else if (method_name == "mfma") method = PGW_METHOD_MFMA;
        else if (method_name != "auto") throw std::runtime_error("--method must be auto, plain or mfma");
        if (save_after != 0) throw std::runtime_error("intermediate saves are not supported: use -s 0");
        if (num_v < 1 || num_v > 16) throw std::runtime_error("-v must be in 1..16");

        std::cout << "Loading instance JSON: " << instance_path << std::endl;
        Instance inst = load_instance(instance_path);
        const std::string op_name = get(a, "--operator", inst.synthetic ? "synthetic" : "eliminated");
        // the deformation variants (cuda/cuprym_deformation.cu:3091-3123, 2578-2623)
        const bool deform = op_name == "augmented" || op_name == "replacement";
        if (inst.synthetic ? op_name != "synthetic" : (op_name != "eliminated" && !deform)) {
            std::cerr << "--operator must be " << (inst.synthetic ? "synthetic" : "one of: eliminated, augmented, replacement")
                      << " for this instance." << std::endl;
            return 1;
        }
        if (deform && !has(a, "--dense-file")) throw std::runtime_error("deformation operators require --dense-file");
        if (op_name == "replacement" && !has(a, "--drop-columns"))
            throw std::runtime_error("replacement operator requires --drop-columns");
        if (op_name != "replacement" && has(a, "--drop-columns"))
            throw std::runtime_error("--drop-columns is only meaningful for replacement");
        if (!is_prime(inst.modulus)) {
            std::cerr << "Modulus " << inst.modulus << " is not prime." << std::endl;
            return 1;
        }
        const uint32_t r = inst.synthetic ? 1 : inst.order;
        if (sector >= r) {
            std::cerr << "--sector must be in 0.." << r - 1 << std::endl;
            return 1;
        }

        pgw_operator_desc desc = make_desc(inst, sector, num_v, method);
        uint64_t rows = 0, cols = 0;
        check(pgw_operator_dims(&desc, &rows, &cols));
        if (!inst.synthetic && sector < inst.expected_sector_columns.size() &&
            (inst.expected_sector_columns[sector] != cols || inst.expected_sector_rows[sector] != rows))
            throw std::runtime_error("eliminated sector dimensions do not match fixture metadata");
        DenseBlock dense;
        if (deform) {
            dense = load_dense_block(get(a, "--dense-file", ""), inst.modulus);
            if (dense.rows != rows || dense.cols != 2)
                throw std::runtime_error("dense block dimensions do not match the eliminated target sector");
            desc.kind = op_name == "augmented" ? PGW_KIND_CYCLIC_AUGMENTED : PGW_KIND_CYCLIC_REPLACEMENT;
            desc.dense_cols = (uint32_t)dense.cols;
            desc.dense_block = dense.values.data();
            if (op_name == "replacement") {
                const std::vector<uint32_t> drops = load_drop_columns(get(a, "--drop-columns", ""), dense.cols, cols);
                desc.drop_cols[0] = drops[0];
                desc.drop_cols[1] = drops[1];
                std::cout << "Dropped base columns: " << drops[0] << " " << drops[1] << std::endl;
            }
            std::cout << "Base F0 dimensions: " << rows << " x " << cols << std::endl;
            check(pgw_operator_dims(&desc, &rows, &cols));
        }
        std::cout << "Sector " << sector << ": " << rows << " x " << cols << " (" << op_name
                  << ", p=" << inst.modulus << ", b=" << num_v << ")" << std::endl;

        pgw_ctx* ctx = nullptr;
        check(pgw_ctx_create(device, &ctx));
        char dev_name[256];
        uint64_t dev_mem = 0;
        check(pgw_ctx_device_info(ctx, dev_name, sizeof(dev_name), &dev_mem));
        std::cout << "Using HIP device " << device << ": " << dev_name << ", global memory "
                  << std::fixed << std::setprecision(2) << dev_mem / 1048576.0 << " MiB" << std::endl;

        // cuda/cuprym_cyclic.cu:2196-2217
        std::vector<uint32_t> dr(rows), dc(cols), v(cols * num_v);
        check(pgw_reference_random_state(seed, sector, inst.modulus, rows, cols, num_v, dr.data(),
                                         dc.data(), v.data()));
        const uint64_t realized = pgw_rowmix_realized_seed(rowmix_seed, op_name.c_str(), sector, rows,
                                                           inst.modulus);
        const uint64_t pc = rows / 2;
        std::vector<pgw_rowmix_pair> pairs((uint64_t)rounds * pc + 1);
        check(pgw_rowmix_plan(realized, rows, rounds, inst.modulus, pairs.data()));
        const uint64_t plan_hash = pgw_rowmix_plan_hash(rows, rounds, rowmix_seed, realized, pairs.data());
        const std::string wdm = sector_filename(out_dir, op_name, sector, num_v, rounds, rowmix_seed);
        if (!no_save) {
            if (std::filesystem::exists(wdm))
                throw std::runtime_error("refusing to overwrite existing WDM file " + wdm);
            if (std::filesystem::exists(rowmix_sidecar_filename(wdm)))
                throw std::runtime_error("refusing to reuse existing rowmix sidecar " + rowmix_sidecar_filename(wdm));
        }
        std::cout << "Row mixing: rounds=" << rounds << ", seed=" << rowmix_seed
                  << ", realized_seed=" << realized << ", plan_hash=" << hex_u64(plan_hash) << std::endl;
        if (!no_save) {
            write_rowmix_sidecar(wdm, op_name, sector, rows, cols, num_v, inst.modulus, seed, rounds,
                                 rowmix_seed, realized, plan_hash, pgw_fnv1a64_u32(dr.data(), rows),
                                 pgw_fnv1a64_u32(dc.data(), cols),
                                 pgw_fnv1a64_u32(v.data(), cols * num_v));
            std::cout << "Wrote rowmix sidecar to " << rowmix_sidecar_filename(wdm) << std::endl;
        }

        pgw_op* op = nullptr;
        check(pgw_operator_create(ctx, &desc, dr.data(), dc.data(), rounds, pairs.data(), &op));
        pgw_op_info info{};
        check(pgw_operator_info(op, &info));
        std::cout << "Operator on device: " << std::setprecision(2) << info.device_bytes / 1048576.0
                  << " MiB, method=" << (info.method == PGW_METHOD_MFMA ? "mfma" : "plain") << std::endl;

        if (has(a, "--validate-rowmix")) {
            // validate_eliminated_cpu_cuda (cuda/cuprym_cyclic.cu:1970-2132): the
            // row-mixing transpose identity on two deterministic blocks, then S X
            // on the device against the exact host apply on three salted blocks,
            // entry by entry; the symmetry <S u, w> = <u, S w> as an extra check
            const uint32_t p = inst.modulus;
            std::vector<uint32_t> x(rows * num_v), y(rows * num_v);
            check(pgw_deterministic_block(rows, num_v, p, 0xA11CE001ULL, x.data()));
            check(pgw_deterministic_block(rows, num_v, p, 0xB0B5EED2ULL, y.data()));
            std::vector<uint32_t> px = x, pty = y;
            host_rowmix(px, pairs, pc, rounds, num_v, p, false);
            host_rowmix(pty, pairs, pc, rounds, num_v, p, true);
            const uint64_t left = dot_mod(px, y, p), right = dot_mod(x, pty, p);
            if (left != right)
                throw std::runtime_error(op_name + " target rowmix transpose identity failed: <Tx,y>=" +
                                         std::to_string(left) + ", <x,T^t y>=" + std::to_string(right));
            std::cout << "Validation matched: " << op_name << " target rowmix transpose identity" << std::endl;
            // host-buffer block-apply on the device: the recovery entry with
            // generators (G_0, G_1) = (0, I) returns D = S u
            std::vector<uint32_t> gens2(2 * num_v * num_v, 0), prod(rows * num_v);
            for (uint32_t c = 0; c < num_v; ++c) gens2[num_v * num_v + c * num_v + c] = 1;
            auto device_apply = [&](const std::vector<uint32_t>& in, std::vector<uint32_t>& out) {
                check(pgw_kernel_vectors(op, in.data(), gens2.data(), 2, out.data(), prod.data()));
            };
            const uint64_t salts[3] = {0x404ULL, 0x505ULL, 0x606ULL};
            std::vector<std::vector<uint32_t>> sx(3);
            for (int idx = 0; idx < 3; ++idx) {
                std::vector<uint32_t> in(cols * num_v), cpu(cols * num_v), gpu(cols * num_v);
                check(pgw_deterministic_block(cols, num_v, p, salts[idx], in.data()));
                device_apply(in, gpu);
                if (inst.synthetic) {  // no host apply for the synthetic operator: symmetry only
                    sx[idx] = std::move(gpu);
                    continue;
                }
                check(pgw_cyclic_apply_cpu(&desc, dr.data(), dc.data(), rounds, pairs.data(), in.data(), num_v,
                                           cpu.data()));
                for (uint64_t q = 0; q < cpu.size(); ++q)
                    if (cpu[q] != gpu[q])
                        throw std::runtime_error(op_name + " square block " + std::to_string(idx) + " mismatch at (" +
                                                 std::to_string(q / num_v) + "," + std::to_string(q % num_v) +
                                                 "): CPU=" + std::to_string(cpu[q]) +
                                                 ", GPU=" + std::to_string(gpu[q]));
                std::cout << "Validation matched: " << op_name << " square block " << idx << " (" << cols << " x "
                          << num_v << ")" << std::endl;
                sx[idx] = std::move(gpu);
            }
            std::vector<uint32_t> u0(cols * num_v), u1(cols * num_v);
            check(pgw_deterministic_block(cols, num_v, p, salts[0], u0.data()));
            check(pgw_deterministic_block(cols, num_v, p, salts[1], u1.data()));
            if (dot_mod(sx[0], u1, p) != dot_mod(u0, sx[1], p))
                throw std::runtime_error("GPU square operator is not symmetric on blocks 0 and 1");
            std::cout << "Validation matched: <S u, w> = <u, S w> on blocks 0 and 1" << std::endl;
        }

        const uint64_t max_len = has(a, "--length") ? get_u64(a, "--length", 0)
                                                     : 2 * cols / num_v + get_u64(a, "--extra-terms", 50);
        const uint64_t L = 1 + 4 * ((max_len - 1 + 3) / 4);
        const uint64_t ut = (uint64_t)num_v * (num_v + 1) / 2;
        std::cout << "Sequence target: " << max_len << " entries (" << L << " after rounding to rounds)"
                  << std::endl;
        std::vector<uint32_t> seq(L * ut), curv(cols * num_v);
        // the WDM text is written while the loop runs: header, Dr, Dc and V
        // at once, the sequence terms as they come back (every 64 rounds)
        std::unique_ptr<pgwio::WdmStream> stream;
        if (!no_save) {
            std::cout << "Streaming WDM state to " << wdm << " during the sequence loop" << std::endl;
            stream.reset(new pgwio::WdmStream(wdm, rows, cols, inst.modulus, dr.data(), dc.data(), v.data(),
                                              num_v, L));
        }
        LoopState loop{std::chrono::steady_clock::now(), stream.get(), seq.data()};
        const auto t0 = loop.start;
        uint64_t got = 0;
        check(pgw_sequence_run(op, v.data(), max_len, seq.data(), curv.data(), &got, progress_cb, &loop));
        auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "\nSector " << sector << " sequence runtime: " << ms << " ms" << std::endl;
        if (ms > 0)
            std::cout << "Sector " << sector << " total throughput: " << got * num_v * 1e3 / ms
                      << "/s, block-applies/s: " << (got - 1) / 2 * 1e3 / ms << std::endl;
        if (no_save) {
            std::cout << "Skipping final WDM save due to --no-save." << std::endl;
        } else {
            if (got != L) throw std::runtime_error("sequence length differs from the streamed WDM header");
            std::cout << "Finishing WDM state " << wdm << "..." << std::endl;
            auto tw = std::chrono::steady_clock::now();
            stream->terms(seq.data(), got);
            stream->finish(curv.data());
            stream.reset();
            std::cout << "Saved " << got << " sequence elements to " << wdm << " ("
                      << std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - tw)
                             .count()
                      << " ms after the loop)" << std::endl;
        }
        check(pgw_operator_destroy(op));
        check(pgw_ctx_destroy(ctx));
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "cuprym_cyclic error: " << e.what() << std::endl;
        return 1;
    }
}

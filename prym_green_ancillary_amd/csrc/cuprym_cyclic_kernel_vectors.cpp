// cuprym_cyclic_kernel_vectors -- kernel-vector recovery from a completed WDM
// file and the generator file of the rank tool (the reference's
// cuprym_cyclic_kernel_vectors, cuda/cuprym_cyclic_kernel_vectors.cu:3010-3203).
//
//   cuprym_cyclic_kernel_vectors INSTANCE -f WDM --sector S [-g GENERATORS]
//       [--preconditioned-output F1] [--original-output F2] [--metadata-output F3]
//       [--operator eliminated|synthetic] [-d DEV] [--rowmix-rounds R] [--rowmix-seed N]
//
// D = sum_k S^k V G_k; certificate F (Dc D) == 0 else exit 1; outputs D and
// Dc D (one vector per line) and a JSON report.

#include <algorithm>
#include <chrono>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

#include "cli_common.h"
#include "wdm.h"

using namespace pgwcli;

namespace {

std::string default_generators(const std::string& wdm) {
    for (const std::string suf : {".wdm.zst", ".wdm"})
        if (wdm.size() >= suf.size() && wdm.compare(wdm.size() - suf.size(), suf.size(), suf) == 0)
            return wdm.substr(0, wdm.size() - suf.size()) + "_generators.txt";
    return wdm + "_generators.txt";
}

// load_generators (cuda/cuprym_cyclic_kernel_vectors.cu:1287-1336): b*b signed
// integers per non-blank line, normalised mod p
std::vector<uint32_t> load_generators(const std::string& path, uint32_t b, uint32_t p) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("failed to open generators file " + path);
    std::vector<uint32_t> out;
    std::string line;
    size_t lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        if (line.find_first_not_of(" \t\r\n") == std::string::npos) continue;
        std::istringstream ls(line);
        for (uint32_t i = 0; i < b * b; ++i) {
            long long v;
            if (!(ls >> v)) throw std::runtime_error("generator line " + std::to_string(lineno) + " has too few entries");
            long long r = v % (long long)p;
            out.push_back((uint32_t)(r < 0 ? r + p : r));
        }
        std::string extra;
        if (ls >> extra)
            throw std::runtime_error("generator line " + std::to_string(lineno) + " has more than num_v*num_v entries");
    }
    if (out.empty()) throw std::runtime_error("generator file does not contain any generator matrices");
    return out;
}

uint32_t powmod(uint64_t a, uint64_t e, uint32_t p) {
    uint64_t r = 1;
    a %= p;
    while (e) {
        if (e & 1) r = r * a % p;
        a = a * a % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

// rank_dense_candidate_columns (cuda/cuprym_cyclic_kernel_vectors.cu:1490-1546)
size_t rank_columns(std::vector<uint32_t> m, size_t rows, size_t cols, uint32_t p,
                    std::vector<size_t>& pivots) {
    pivots.clear();
    size_t rank = 0;
    for (size_t col = 0; col < cols && rank < rows; ++col) {
        size_t piv = rows;
        for (size_t r = rank; r < rows; ++r)
            if (m[r * cols + col] % p) { piv = r; break; }
        if (piv == rows) continue;
        if (piv != rank)
            for (size_t c = 0; c < cols; ++c) std::swap(m[rank * cols + c], m[piv * cols + c]);
        const uint32_t inv = powmod(m[rank * cols + col], p - 2, p);
        for (size_t c = col; c < cols; ++c) m[rank * cols + c] = (uint32_t)((uint64_t)m[rank * cols + c] * inv % p);
        for (size_t r = 0; r < rows; ++r) {
            if (r == rank) continue;
            const uint64_t f = m[r * cols + col] % p;
            if (!f) continue;
            for (size_t c = col; c < cols; ++c) {
                const uint64_t sub = f * m[rank * cols + c] % p;
                m[r * cols + c] = (uint32_t)((m[r * cols + c] % p + p - sub) % p);
            }
        }
        pivots.push_back(col);
        ++rank;
    }
    return rank;
}

void save_vectors(const std::string& path, const std::vector<uint32_t>& d, size_t rows, size_t b,
                  const std::vector<uint32_t>* scale, uint32_t p) {
    std::filesystem::path out(path);
    if (out.has_parent_path()) std::filesystem::create_directories(out.parent_path());
    const std::string tmp = path + ".tmp";
    {
        std::ofstream f(tmp);
        if (!f) throw std::runtime_error("failed to open output file " + tmp);
        for (size_t c = 0; c < b; ++c) {
            for (size_t r = 0; r < rows; ++r) {
                uint64_t v = d[r * b + c] % p;
                if (scale) v = v * (*scale)[r] % p;
                if (r) f << ' ';
                f << v;
            }
            f << '\n';
        }
        if (!f) throw std::runtime_error("failed while writing output file " + tmp);
    }
    std::filesystem::rename(tmp, path);
}

std::vector<size_t> column_nnz(const std::vector<uint32_t>& m, size_t rows, size_t b, uint32_t p) {
    std::vector<size_t> nnz(b, 0);
    for (size_t r = 0; r < rows; ++r)
        for (size_t c = 0; c < b; ++c)
            if (m[r * b + c] % p) ++nnz[c];
    return nnz;
}

std::string list(const std::vector<size_t>& v) {
    std::ostringstream o;
    for (size_t i = 0; i < v.size(); ++i) o << (i ? ", " : "") << v[i];
    return o.str();
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const std::vector<OptSpec> specs = {
            {{"-f", "--wdm-file"}, false}, {{"-g", "--generators-file"}, false},
            {{"--preconditioned-output"}, false}, {{"--original-output"}, false},
            {{"--metadata-output"}, false}, {{"--operator"}, false}, {{"--sector"}, false},
            {{"-d", "--device"}, false}, {{"--rowmix-rounds"}, false}, {{"--rowmix-seed"}, false},
            {{"--validate-rowmix"}, true}, {{"--method"}, false}, {{"-h", "--help"}, true},
        };
        Args a = parse_args(argc, argv, specs);
        if (has(a, "--help") || a.positional.size() != 1 || !has(a, "--wdm-file") || !has(a, "--sector")) {
            std::cerr << "usage: cuprym_cyclic_kernel_vectors INSTANCE -f WDM --sector S [-g GEN] "
                         "[--preconditioned-output F] [--original-output F] [--metadata-output F] "
                         "[-d DEV] [--rowmix-rounds R] [--rowmix-seed N]\n";
            return has(a, "--help") ? 0 : 1;
        }
        const std::string instance_path = a.positional[0];
        const std::string wdm = get(a, "--wdm-file", "");
        const std::string gen_path = get(a, "--generators-file", default_generators(wdm));
        const std::string out1 = get(a, "--preconditioned-output", wdm + "_nullvectors_1.txt");
        const std::string out2 = get(a, "--original-output", wdm + "_nullvectors_2.txt");
        const std::string meta = get(a, "--metadata-output", wdm + "_nullvectors_report.json");
        const uint32_t sector = (uint32_t)get_u64(a, "--sector", 0);
        for (const auto& f : {instance_path, wdm, gen_path})
            if (!std::filesystem::exists(f)) throw std::runtime_error("file not found: " + f);

        Instance inst = load_instance(instance_path);
        const std::string op_name = get(a, "--operator", inst.synthetic ? "synthetic" : "eliminated");
        if (!is_prime(inst.modulus)) throw std::runtime_error("modulus is not prime");
        if (sector >= (inst.synthetic ? 1u : inst.order)) throw std::runtime_error("--sector outside 0..r-1");
        pgwio::WdmInitial w = pgwio::load_wdm_initial(wdm);
        if (w.prime != inst.modulus) throw std::runtime_error("WDM prime does not match the instance modulus");
        const uint32_t b = w.block;
        pgw_operator_desc desc = make_desc(inst, sector, b, PGW_METHOD_AUTO);
        const std::string m = get(a, "--method", "auto");
        desc.method = m == "plain" ? PGW_METHOD_PLAIN : (m == "mfma" ? PGW_METHOD_MFMA : PGW_METHOD_AUTO);
        uint64_t rows = 0, cols = 0;
        check(pgw_operator_dims(&desc, &rows, &cols));
        if (rows != w.rows || cols != w.cols) throw std::runtime_error("WDM dimensions do not match the operator");

        // resolve_rowmix_options (cuprym_cyclic_kernel_vectors.cu:1554-1643): the
        // sidecar wins; its hashes must match the WDM's preconditioners and block
        Sidecar sc = read_rowmix_sidecar(wdm);
        uint32_t rounds = (uint32_t)get_u64(a, "--rowmix-rounds", 2);
        uint64_t rseed = get_u64(a, "--rowmix-seed", 10001);
        if (sc.present) {
            rounds = sc.rounds;
            rseed = sc.seed;
            if (sc.dr_hash != pgw_fnv1a64_u32(w.dr.data(), rows) ||
                sc.dc_hash != pgw_fnv1a64_u32(w.dc.data(), cols) ||
                sc.v_hash != pgw_fnv1a64_u32(w.v.data(), cols * b))
                throw std::runtime_error("rowmix sidecar hashes do not match the WDM initial state");
        }
        const uint64_t realized = pgw_rowmix_realized_seed(rseed, op_name.c_str(), sector, rows, inst.modulus);
        std::vector<pgw_rowmix_pair> pairs((uint64_t)rounds * (rows / 2) + 1);
        check(pgw_rowmix_plan(realized, rows, rounds, inst.modulus, pairs.data()));
        const uint64_t plan_hash = pgw_rowmix_plan_hash(rows, rounds, rseed, realized, pairs.data());
        if (sc.present && (sc.realized_seed != realized || sc.plan_hash != plan_hash))
            throw std::runtime_error("rowmix sidecar plan does not match the regenerated plan");

        std::vector<uint32_t> gens = load_generators(gen_path, b, inst.modulus);
        const uint64_t ngen = gens.size() / ((uint64_t)b * b);
        std::cout << "Recovery dimensions: rows=" << rows << ", columns=" << cols << ", v=" << b
                  << ", WDM sequence length=" << w.length << ", generator matrices=" << ngen << std::endl;

        pgw_ctx* ctx = nullptr;
        check(pgw_ctx_create((int)get_u64(a, "--device", 0), &ctx));
        pgw_op* op = nullptr;
        check(pgw_operator_create(ctx, &desc, w.dr.data(), w.dc.data(), rounds, pairs.data(), &op));
        std::vector<uint32_t> d(cols * b), prod(rows * b);
        auto t0 = std::chrono::steady_clock::now();
        check(pgw_kernel_vectors(op, w.v.data(), gens.data(), ngen, d.data(), prod.data()));
        const long long ms = std::chrono::duration_cast<std::chrono::milliseconds>(
                                 std::chrono::steady_clock::now() - t0).count();
        check(pgw_operator_destroy(op));
        check(pgw_ctx_destroy(ctx));

        const std::vector<size_t> prod_nnz = column_nnz(prod, rows, b, inst.modulus);
        std::cout << "A*(source-preconditioned candidate) nonzeros by vector: " << list(prod_nnz) << std::endl;
        size_t total = 0;
        for (size_t x : prod_nnz) total += x;
        if (total) throw std::runtime_error("kernel-vector certificate failed");
        std::cout << "Kernel-vector certificate passed." << std::endl;
        const std::vector<size_t> cand_nnz = column_nnz(d, cols, b, inst.modulus);
        size_t ctotal = 0;
        for (size_t x : cand_nnz) ctotal += x;
        if (!ctotal) throw std::runtime_error("all candidate vectors are zero; no kernel vector was produced");
        std::vector<size_t> pivots;
        const size_t rank = rank_columns(d, cols, b, inst.modulus, pivots);
        std::cout << "Candidate block rank: " << rank << "/" << b << std::endl;
        save_vectors(out1, d, cols, b, nullptr, inst.modulus);
        save_vectors(out2, d, cols, b, &w.dc, inst.modulus);
        {
            const std::string tmp = meta + ".tmp";
            std::ofstream o(tmp);
            o << "{\n  \"type\": \"prym_cyclic_kernel_vectors_report\",\n  \"version\": 1,\n";
            o << "  \"instance\": \"" << instance_path << "\",\n  \"wdm_file\": \"" << wdm << "\",\n";
            o << "  \"generators_file\": \"" << gen_path << "\",\n  \"operator\": \"" << op_name << "\",\n";
            o << "  \"sector\": " << sector << ",\n  \"rows\": " << rows << ",\n  \"columns\": " << cols << ",\n";
            o << "  \"prime\": " << inst.modulus << ",\n  \"vector_count\": " << b << ",\n";
            o << "  \"wdm_sequence_length\": " << w.length << ",\n  \"generator_count\": " << ngen << ",\n";
            o << "  \"runtime_ms\": " << ms << ",\n  \"rowmix\": {\n    \"rounds\": " << rounds << ",\n";
            o << "    \"seed\": " << rseed << ",\n    \"realized_seed\": " << realized << ",\n";
            o << "    \"hash_fnv1a64\": \"" << hex_u64(plan_hash) << "\"\n  },\n";
            o << "  \"candidate_nonzeros\": [" << list(cand_nnz) << "],\n";
            o << "  \"product_nonzeros\": [" << list(prod_nnz) << "],\n";
            o << "  \"candidate_rank\": " << rank << ",\n  \"pivot_columns\": [" << list(pivots) << "]\n}\n";
            o.close();
            if (!o) throw std::runtime_error("failed while writing metadata output " + tmp);
            std::filesystem::rename(tmp, meta);
        }
        std::cout << "Saved preconditioned candidates to " << out1 << "\nSaved original-coordinate candidates to "
                  << out2 << "\nSaved metadata to " << meta << std::endl;
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "cuprym_cyclic_kernel_vectors error: " << e.what() << std::endl;
        return 1;
    }
}

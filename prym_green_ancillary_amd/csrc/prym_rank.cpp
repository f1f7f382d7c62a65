// prym_rank.cpp -- `prym-rank FILE.wdm[.zst] [-g] [--threads T]`: the rank
// lower bound (and with -g the matrix generator) from a completed WDM file.
// Command line, messages and output files of the BCW tool
// (third_party/bcw/src/main.rs:17-139):
//   <stem>_result.txt      "Matrix size: R x C\nRank: r\n"
//   <stem>_generators.txt  one b x b matrix per line, row-major, "v " per value
// where <stem> is FILE without ".wdm.zst" / ".wdm".  Errors print
// "Error: <message>" on stderr and exit 1, writing no result.

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pgw.h"
#include "cli_common.h"
#include "wdm.h"

namespace {

// output_path_with_suffix (main.rs:124-131)
std::string output_path_with_suffix(const std::string& path, const std::string& suffix) {
    auto strip = [&](const std::string& ext) {
        return path.size() >= ext.size() && path.compare(path.size() - ext.size(), ext.size(), ext) == 0;
    };
    std::string stem = path;
    if (strip(".wdm.zst")) stem = path.substr(0, path.size() - 8);
    else if (strip(".wdm")) stem = path.substr(0, path.size() - 4);
    return stem + suffix;
}

struct RankHandle {
    pgw_rank* r = nullptr;
    ~RankHandle() { pgw_rank_destroy(r); }
};

void write_text(const std::string& path, const std::string& text) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot create " + path);
    const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
    if (std::fclose(f) != 0 || !ok) throw std::runtime_error("cannot write " + path);
}

}  // namespace

int main(int argc, char** argv) {
    using namespace pgwcli;
    try {
        const std::vector<OptSpec> specs = {
            {{"-g", "--generator"}, true},
            {{"-t", "--threads"}, false},
            {{"-h", "--help"}, true},
        };
        Args a = parse_args(argc, argv, specs);
        if (has(a, "--help") || a.positional.size() != 1) {
            std::fprintf(has(a, "--help") ? stdout : stderr,
                         "usage: prym-rank FILE.wdm[.zst] [-g|--generator] [-t|--threads T]\n");
            return has(a, "--help") ? 0 : 1;
        }
        const uint64_t threads = get_u64(a, "--threads", 4);
        if (threads == 0) throw std::runtime_error("--threads must be positive");
        const std::string path = a.positional[0];

        std::printf("Loading completed WDM file %s...\n", path.c_str());
        const pgwio::WdmFull w = pgwio::load_wdm_full(path);
        const uint64_t usable = w.length ? w.length - 1 : 0;
        std::printf("File data: prime %u, matrix size %llux%llu, num_v %u, usable sequence entries %llu.\n",
                    w.prime, (unsigned long long)w.rows, (unsigned long long)w.cols, w.block,
                    (unsigned long long)usable);
        RankHandle h;
        check(pgw_rank_create(w.prime, w.block, w.rows, w.cols, PGW_RANK_DEFAULT_THRESHOLD,
                              PGW_RANK_DEFAULT_MAX_NLEN_BUFFER, (uint32_t)threads, &h.r));
        if (usable == 0)
            throw std::runtime_error(
                "completed WDM file contains no usable sequence terms after removing the initial Gram term");
        pgw_rank_info info;
        check(pgw_rank_step(h.r, w.seq.data(), w.length, &info));
        std::printf("\nProcessed %llu sequence entries; delta_spread=%llu, rank_lower_bound=%llu.\n",
                    (unsigned long long)info.processed, (unsigned long long)info.delta_spread,
                    (unsigned long long)info.rank);
        const uint64_t min_dim = w.rows < w.cols ? w.rows : w.cols;
        if (info.rank > min_dim || info.status == PGW_RANK_EXCEEDS_DIMENSION)
            throw std::runtime_error("computed rank lower bound exceeds the matrix dimensions");
        if (info.status != PGW_RANK_SUCCESS)
            throw std::runtime_error(
                "completed WDM sequence is insufficient: processed " + std::to_string(info.processed) +
                " terms, delta spread " + std::to_string(info.delta_spread) + " (required " +
                std::to_string(PGW_RANK_DEFAULT_THRESHOLD) + "); no rank result was written");
        std::printf("Final rank lower bound: %llu\n", (unsigned long long)info.rank);
        if (has(a, "--generator")) {
            const std::string gpath = output_path_with_suffix(path, "_generators.txt");
            std::printf("Writing generator matrices to %s\n", gpath.c_str());
            uint64_t terms = 0;
            check(pgw_rank_generator(h.r, nullptr, 0, &terms));
            const uint64_t bb = (uint64_t)w.block * w.block;
            std::vector<uint32_t> gen(terms * bb);
            check(pgw_rank_generator(h.r, gen.data(), gen.size(), &terms));
            std::string text;
            text.reserve(gen.size() * 11);
            for (uint64_t k = 0; k < terms; ++k) {
                for (uint64_t e = 0; e < bb; ++e) {
                    text += std::to_string(gen[k * bb + e]);
                    text += ' ';
                }
                text += '\n';
            }
            write_text(gpath, text);
        }
        const std::string rpath = output_path_with_suffix(path, "_result.txt");
        write_text(rpath, "Matrix size: " + std::to_string(w.rows) + " x " + std::to_string(w.cols) +
                              "\nRank: " + std::to_string(info.rank) + "\n");
        std::printf("Wrote %s\n", rpath.c_str());
        return 0;
    } catch (const std::exception& e) {
        std::fflush(stdout);
        std::fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
}

// cli_common.cpp -- instance JSON scanner, option parser, file names, sidecar.

#include "cli_common.h"

#include <algorithm>
#include <cctype>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace pgwcli {

namespace {

std::string read_text(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("failed to open " + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

// Value text of "key" inside `text` (the first occurrence at any depth), as
// the reference's scanner reads it (cuda/cuprym_cyclic.cu:252-312).
bool find_value(const std::string& text, const std::string& key, size_t& begin) {
    const std::string needle = "\"" + key + "\"";
    size_t pos = 0;
    while ((pos = text.find(needle, pos)) != std::string::npos) {
        size_t q = pos + needle.size();
        while (q < text.size() && std::isspace((unsigned char)text[q])) ++q;
        if (q < text.size() && text[q] == ':') {
            ++q;
            while (q < text.size() && std::isspace((unsigned char)text[q])) ++q;
            begin = q;
            return true;
        }
        pos = q;
    }
    return false;
}

size_t matching(const std::string& t, size_t open) {
    const char o = t[open], c = o == '{' ? '}' : ']';
    int depth = 0;
    bool in_str = false, esc = false;
    for (size_t i = open; i < t.size(); ++i) {
        char ch = t[i];
        if (in_str) {
            if (esc) esc = false;
            else if (ch == '\\') esc = true;
            else if (ch == '"') in_str = false;
            continue;
        }
        if (ch == '"') in_str = true;
        else if (ch == o) ++depth;
        else if (ch == c && --depth == 0) return i;
    }
    throw std::runtime_error("unbalanced JSON");
}

std::string object(const std::string& t, const std::string& key, bool required = true) {
    size_t b;
    if (!find_value(t, key, b) || t[b] != '{') {
        if (required) throw std::runtime_error("missing JSON object '" + key + "'");
        return {};
    }
    return t.substr(b, matching(t, b) - b + 1);
}

std::string array(const std::string& t, const std::string& key, bool required = true) {
    size_t b;
    if (!find_value(t, key, b) || t[b] != '[') {
        if (required) throw std::runtime_error("missing JSON array '" + key + "'");
        return {};
    }
    return t.substr(b, matching(t, b) - b + 1);
}

uint64_t number(const std::string& t, const std::string& key) {
    size_t b;
    if (!find_value(t, key, b) || !std::isdigit((unsigned char)t[b]))
        throw std::runtime_error("missing or non-integer JSON key '" + key + "'");
    uint64_t v = 0;
    for (; b < t.size() && std::isdigit((unsigned char)t[b]); ++b) {
        uint64_t d = (uint64_t)(t[b] - '0');
        if (v > (std::numeric_limits<uint64_t>::max() - d) / 10) throw std::runtime_error(key + " overflows u64");
        v = v * 10 + d;
    }
    return v;
}

// all unsigned integers of an array text, flattened (nested arrays allowed)
std::vector<uint64_t> integers(const std::string& arr) {
    std::vector<uint64_t> out;
    for (size_t i = 0; i < arr.size();) {
        if (std::isdigit((unsigned char)arr[i])) {
            uint64_t v = 0;
            while (i < arr.size() && std::isdigit((unsigned char)arr[i])) v = v * 10 + (uint64_t)(arr[i++] - '0');
            out.push_back(v);
        } else {
            if (arr[i] == '-') throw std::runtime_error("negative entry in an unsigned JSON array");
            ++i;
        }
    }
    return out;
}

template <typename T>
std::vector<T> narrow(const std::vector<uint64_t>& v, uint64_t bound, const std::string& what) {
    std::vector<T> out;
    out.reserve(v.size());
    for (uint64_t x : v) {
        if (x >= bound) throw std::runtime_error(what + " entry " + std::to_string(x) + " is out of range");
        out.push_back((T)x);
    }
    return out;
}

}  // namespace

Instance load_instance(const std::string& path) {
    const std::string text = read_text(path);
    Instance I;
    I.modulus = (uint32_t)number(text, "modulus");
    std::string syn = object(text, "synthetic", false);
    if (!syn.empty()) {
        I.synthetic = true;
        I.n = (uint32_t)number(syn, "n");
        I.k = (uint32_t)number(syn, "k");
        I.m_src = (uint32_t)number(syn, "M");
        I.m_tgt = (uint32_t)number(syn, "M_prime");
        I.table_seed_d = number(syn, "table_seed_d");
        I.table_seed_e = number(syn, "table_seed_e");
        const uint64_t count = (uint64_t)I.n * I.m_tgt * I.m_src;
        I.table_d.resize(count);
        I.table_e.resize(count);
        check(pgw_synthetic_table(I.table_seed_d, I.modulus, count, I.table_d.data()));
        check(pgw_synthetic_table(I.table_seed_e, I.modulus, count, I.table_e.data()));
        return I;
    }
    // prym-cyclic-instance-v1 fields (cuda/cuprym_cyclic.cu:483-609)
    I.genus = (uint32_t)number(text, "genus");
    const std::string cyc = object(text, "cyclic");
    I.order = (uint32_t)number(cyc, "order");
    if (I.genus < 6 || I.genus % 2) throw std::runtime_error("cyclic Prym path expects an even genus >= 6");
    const uint32_t n = I.genus - 3;
    const std::string w = object(text, "weights");
    I.weight_v = narrow<uint16_t>(integers(array(w, "V")), I.order, "weights.V");
    I.weight_a0 = narrow<uint16_t>(integers(array(w, "A0")), I.order, "weights.A0");
    I.weight_a1 = narrow<uint16_t>(integers(array(w, "A1")), I.order, "weights.A1");
    if (I.weight_v.size() != n || I.weight_a0.size() != I.genus || I.weight_a1.size() != n)
        throw std::runtime_error("cyclic weight vector lengths do not match genus metadata");
    const std::string mu = object(object(text, "artinian"), "mu");
    if (number(mu, "n") != n || number(mu, "g") != I.genus)
        throw std::runtime_error("artinian.mu dimensions do not match cyclic genus metadata");
    I.mu = narrow<uint32_t>(integers(array(mu, "values")), I.modulus, "mu.values");
    if (I.mu.size() != (uint64_t)n * n * I.genus)
        throw std::runtime_error("artinian.mu dimensions do not match cyclic genus metadata");
    const std::string e = object(text, "third_section_elimination", false);
    if (!e.empty()) {
        I.has_elimination = true;
        I.w_index = (uint32_t)number(e, "w_V_index");
        I.w_weight = (uint32_t)number(e, "w_weight");
        I.kernel_weights = narrow<uint16_t>(integers(array(e, "kernel_weights")), I.order, "kernel_weights");
        I.right_inverse = narrow<uint32_t>(integers(array(e, "right_inverse_A0_by_A1")), I.modulus,
                                           "right_inverse_A0_by_A1");
        I.kernel_basis = narrow<uint32_t>(integers(array(e, "kernel_basis_A0_by_3")), I.modulus,
                                          "kernel_basis_A0_by_3");
        if (I.right_inverse.size() != (uint64_t)I.genus * n)
            throw std::runtime_error("elimination right inverse has wrong shape");
        if (I.kernel_basis.size() != (uint64_t)I.genus * I.kernel_weights.size())
            throw std::runtime_error("elimination kernel basis has wrong shape");
        I.expected_sector_columns = integers(array(e, "sector_columns", false));
        I.expected_sector_rows = integers(array(e, "sector_rows", false));
    }
    return I;
}

DenseBlock load_dense_block(const std::string& path, uint32_t prime) {
    const std::string text = read_text(path);
    DenseBlock D;
    D.rows = number(text, "rows");
    D.cols = number(text, "columns");
    if (number(text, "prime") != prime) throw std::runtime_error("dense block prime does not match instance modulus");
    std::vector<uint64_t> raw = integers(array(text, "values"));
    if (raw.size() != D.rows * D.cols) throw std::runtime_error("dense block has the wrong number of values");
    D.values = narrow<uint32_t>(raw, prime, "dense block");
    return D;
}

std::vector<uint32_t> load_drop_columns(const std::string& path, size_t expected, uint64_t base_cols) {
    const std::string text = read_text(path);
    std::vector<uint64_t> raw = integers(array(text, "indices"));
    if (raw.size() != expected) throw std::runtime_error("drop-column file has the wrong number of indices");
    std::vector<uint32_t> out = narrow<uint32_t>(raw, base_cols, "drop-column index (outside the base column range)");
    std::sort(out.begin(), out.end());
    for (size_t i = 1; i < out.size(); ++i)
        if (out[i] == out[i - 1]) throw std::runtime_error("drop-column indices must be distinct");
    return out;
}

pgw_operator_desc make_desc(Instance& I, uint32_t sector, uint32_t block, uint32_t method) {
    pgw_operator_desc d{};
    d.prime = I.modulus;
    d.block = block;
    d.method = method;
    if (I.synthetic) {
        d.kind = PGW_KIND_SYNTHETIC;
        d.n = I.n;
        d.k = I.k;
        d.m_src = I.m_src;
        d.m_tgt = I.m_tgt;
        d.table_d = I.table_d.data();
        d.table_e = I.table_e.data();
    } else {
        if (!I.has_elimination) throw std::runtime_error("instance does not contain third_section_elimination data");
        d.kind = PGW_KIND_CYCLIC_ELIMINATED;
        d.genus = I.genus;
        d.order = I.order;
        d.sector = sector;
        d.mu = I.mu.data();
        d.weight_v = I.weight_v.data();
        d.weight_a0 = I.weight_a0.data();
        d.weight_a1 = I.weight_a1.data();
        d.w_index = I.w_index;
        d.w_weight = I.w_weight;
        d.kernel_dim = (uint32_t)I.kernel_weights.size();
        d.kernel_weights = I.kernel_weights.data();
        d.right_inverse = I.right_inverse.data();
        d.kernel_basis = I.kernel_basis.data();
    }
    return d;
}

Args parse_args(int argc, char** argv, const std::vector<OptSpec>& specs) {
    Args a;
    for (int i = 1; i < argc; ++i) {
        std::string s = argv[i];
        std::string value;
        bool has_inline = false;
        if (s.rfind("--", 0) == 0 && s.find('=') != std::string::npos) {
            value = s.substr(s.find('=') + 1);
            s = s.substr(0, s.find('='));
            has_inline = true;
        }
        if (s.size() > 1 && s[0] == '-' && !std::isdigit((unsigned char)s[1])) {
            const OptSpec* hit = nullptr;
            for (const auto& sp : specs)
                for (const auto& nm : sp.names)
                    if (nm == s) hit = &sp;
            if (!hit) throw std::runtime_error("unknown option " + s);
            const std::string canonical = hit->names.back();
            if (hit->is_flag) {
                a.flags[canonical] = true;
            } else {
                if (!has_inline) {
                    if (i + 1 >= argc) throw std::runtime_error("option " + s + " needs a value");
                    value = argv[++i];
                }
                a.values[canonical] = value;
            }
        } else {
            a.positional.push_back(s);
        }
    }
    return a;
}

bool has(const Args& a, const std::string& c) { return a.flags.count(c) || a.values.count(c); }

std::string get(const Args& a, const std::string& c, const std::string& dflt) {
    auto it = a.values.find(c);
    return it == a.values.end() ? dflt : it->second;
}

uint64_t get_u64(const Args& a, const std::string& c, uint64_t dflt) {
    auto it = a.values.find(c);
    if (it == a.values.end()) return dflt;
    const std::string& s = it->second;
    if (s.empty() || s.find_first_not_of("0123456789") != std::string::npos)
        throw std::runtime_error("option " + c + " needs a nonnegative integer, got '" + s + "'");
    return std::stoull(s);
}

std::string sector_filename(const std::string& out_dir, const std::string& op, uint32_t sector,
                            uint32_t num_v, uint32_t rounds, uint64_t rowmix_seed) {
    // cuda/cuprym_cyclic.cu:1295-1312
    std::filesystem::path dir(out_dir.empty() ? "." : out_dir);
    std::filesystem::create_directories(dir);
    std::ostringstream name;
    name << op << "-sector" << std::setw(2) << std::setfill('0') << sector << "-v" << num_v;
    if (rounds > 0) name << "-rowmix-r" << rounds << "-s" << rowmix_seed;
    name << ".wdm.zst";
    return (dir / name.str()).string();
}

std::string rowmix_sidecar_filename(const std::string& wdm) {
    const std::string suf = ".wdm.zst";
    if (wdm.size() >= suf.size() && wdm.compare(wdm.size() - suf.size(), suf.size(), suf) == 0)
        return wdm.substr(0, wdm.size() - suf.size()) + ".rowmix.json";
    return wdm + ".rowmix.json";
}

std::string hex_u64(uint64_t v) {
    std::ostringstream out;
    out << "0x" << std::hex << std::setw(16) << std::setfill('0') << v;
    return out.str();
}

void write_rowmix_sidecar(const std::string& wdm, const std::string& op, uint32_t sector,
                          uint64_t rows, uint64_t cols, uint32_t num_v, uint32_t prime,
                          uint64_t numerical_seed, uint32_t rounds, uint64_t rowmix_seed,
                          uint64_t realized_seed, uint64_t plan_hash, uint64_t dr_hash,
                          uint64_t dc_hash, uint64_t v_hash) {
    // same keys and layout as write_rowmix_sidecar (cuda/cuprym_cyclic.cu:1336-1400)
    const std::string path = rowmix_sidecar_filename(wdm), tmp = path + ".tmp";
    std::ofstream out(tmp, std::ios::binary);
    if (!out) throw std::runtime_error("failed to open rowmix sidecar " + tmp);
    out << "{\n";
    out << "  \"type\": \"prym_cyclic_rowmix_sidecar\",\n";
    out << "  \"version\": 1,\n";
    out << "  \"algorithm_version\": \"paired-sl2-target-v1\",\n";
    out << "  \"wdm_file\": \"" << wdm << "\",\n";
    out << "  \"operator\": \"" << op << "\",\n";
    out << "  \"sector\": " << sector << ",\n";
    out << "  \"rows\": " << rows << ",\n";
    out << "  \"columns\": " << cols << ",\n";
    out << "  \"vector_count\": " << num_v << ",\n";
    out << "  \"prime\": " << prime << ",\n";
    out << "  \"numerical_seed\": " << numerical_seed << ",\n";
    out << "  \"rowmix\": {\n";
    out << "    \"rounds\": " << rounds << ",\n";
    out << "    \"seed\": " << rowmix_seed << ",\n";
    out << "    \"realized_seed\": " << realized_seed << ",\n";
    out << "    \"rng\": \"splitmix64-v1\",\n";
    out << "    \"target_rows\": " << rows << ",\n";
    out << "    \"pair_count_per_round\": [";
    for (uint32_t i = 0; i < rounds; ++i) out << (i ? ", " : "") << rows / 2;
    out << "],\n";
    out << "    \"recipe\": \"Starting from realized_seed, for each round reset permutation to the "
           "previous round's final order, run descending Fisher-Yates using splitmix64 "
           "rejection-bounded draws, pair consecutive coordinates, then draw nonzero a,b in F_p "
           "by rejection-bounded draws in 1..p-1.\",\n";
    out << "    \"hash_fnv1a64\": \"" << hex_u64(plan_hash) << "\"\n";
    out << "  },\n";
    out << "  \"hashes\": {\n";
    out << "    \"row_preconditioner_fnv1a64\": \"" << hex_u64(dr_hash) << "\",\n";
    out << "    \"column_preconditioner_fnv1a64\": \"" << hex_u64(dc_hash) << "\",\n";
    out << "    \"initial_block_fnv1a64\": \"" << hex_u64(v_hash) << "\"\n";
    out << "  }\n";
    out << "}\n";
    out.close();
    if (!out) throw std::runtime_error("failed while writing " + tmp);
    std::filesystem::rename(tmp, path);
}

Sidecar read_rowmix_sidecar(const std::string& wdm) {
    Sidecar s;
    const std::string path = rowmix_sidecar_filename(wdm);
    if (!std::filesystem::exists(path)) return s;
    const std::string text = read_text(path);
    s.present = true;
    const std::string rm = object(text, "rowmix");
    s.rounds = (uint32_t)number(rm, "rounds");
    s.seed = number(rm, "seed");
    s.realized_seed = number(rm, "realized_seed");
    auto hex_of = [&](const std::string& t, const std::string& key) {
        size_t b;
        if (!find_value(t, key, b) || t[b] != '"') throw std::runtime_error("sidecar key " + key + " missing");
        return std::stoull(t.substr(b + 1, t.find('"', b + 1) - b - 1), nullptr, 16);
    };
    s.plan_hash = hex_of(rm, "hash_fnv1a64");
    const std::string h = object(text, "hashes");
    s.dr_hash = hex_of(h, "row_preconditioner_fnv1a64");
    s.dc_hash = hex_of(h, "column_preconditioner_fnv1a64");
    s.v_hash = hex_of(h, "initial_block_fnv1a64");
    return s;
}

void check(int rc) {
    if (rc != PGW_OK) throw std::runtime_error(pgw_last_error());
}

bool is_prime(uint32_t p) {
    if (p < 2) return false;
    for (uint32_t q = 2; (uint64_t)q * q <= p; ++q)
        if (p % q == 0) return false;
    return true;
}

}  // namespace pgwcli

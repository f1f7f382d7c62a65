// cli_common.h -- shared pieces of the command-line drivers: instance JSON,
// option parsing, sector file names and the row-mix sidecar.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/pgw.h"

namespace pgwcli {

// Instance spec: either the reference's prym-cyclic JSON (genus, modulus,
// cyclic.order, weights, artinian.mu, third_section_elimination) or a synthetic
// sector spec {"modulus", "synthetic": {n, k, M, M_prime, table_seed_d, table_seed_e}}.
struct Instance {
    bool synthetic = false;
    uint32_t modulus = 0;
    // cyclic
    uint32_t genus = 0, order = 1;
    std::vector<uint32_t> mu;
    std::vector<uint16_t> weight_v, weight_a0, weight_a1;
    bool has_elimination = false;
    uint32_t w_index = 0, w_weight = 0;
    std::vector<uint16_t> kernel_weights;
    std::vector<uint32_t> right_inverse, kernel_basis;
    std::vector<uint64_t> expected_sector_columns, expected_sector_rows;
    // synthetic
    uint32_t n = 0, k = 0, m_src = 0, m_tgt = 0;
    uint64_t table_seed_d = 0, table_seed_e = 0;
    std::vector<uint32_t> table_d, table_e;
};

Instance load_instance(const std::string& path);

// Dense deformation block file {"rows", "columns", "prime", "values"}, values
// row-major and reduced (load_dense_block_file, cuda/cuprym_deformation.cu:426-450).
struct DenseBlock {
    uint64_t rows = 0, cols = 0;
    std::vector<uint32_t> values;
};
DenseBlock load_dense_block(const std::string& path, uint32_t prime);
// Drop-column file {"indices": [...]}: `expected` distinct indices below
// `base_cols`, returned ascending (load_drop_columns_file, :452-472).
std::vector<uint32_t> load_drop_columns(const std::string& path, size_t expected, uint64_t base_cols);
// Fills a pgw_operator_desc pointing into `inst` (which must outlive it).
pgw_operator_desc make_desc(Instance& inst, uint32_t sector, uint32_t block, uint32_t method);

// Minimal option parser: positional arguments plus --name value / -x value
// / flags; unknown options are errors.
struct Args {
    std::vector<std::string> positional;
    std::map<std::string, std::string> values;
    std::map<std::string, bool> flags;
};
struct OptSpec {
    std::vector<std::string> names;  // e.g. {"-d", "--device"}
    bool is_flag;
};
Args parse_args(int argc, char** argv, const std::vector<OptSpec>& specs);
bool has(const Args& a, const std::string& canonical);
std::string get(const Args& a, const std::string& canonical, const std::string& dflt);
uint64_t get_u64(const Args& a, const std::string& canonical, uint64_t dflt);

std::string sector_filename(const std::string& out_dir, const std::string& op, uint32_t sector,
                            uint32_t num_v, uint32_t rounds, uint64_t rowmix_seed);
std::string rowmix_sidecar_filename(const std::string& wdm);
std::string hex_u64(uint64_t v);
void write_rowmix_sidecar(const std::string& wdm, const std::string& op, uint32_t sector,
                          uint64_t rows, uint64_t cols, uint32_t num_v, uint32_t prime,
                          uint64_t numerical_seed, uint32_t rounds, uint64_t rowmix_seed,
                          uint64_t realized_seed, uint64_t plan_hash, uint64_t dr_hash,
                          uint64_t dc_hash, uint64_t v_hash);
// Reads the sidecar's rowmix rounds / seed / plan hash (cuprym_cyclic_kernel_vectors.cu:1554-1643).
struct Sidecar {
    bool present = false;
    uint32_t rounds = 0;
    uint64_t seed = 0, realized_seed = 0, plan_hash = 0;
    uint64_t dr_hash = 0, dc_hash = 0, v_hash = 0;
};
Sidecar read_rowmix_sidecar(const std::string& wdm);

void check(int rc);  // throws std::runtime_error with pgw_last_error() on failure
bool is_prime(uint32_t p);

}  // namespace pgwcli

// pgw_host.cpp -- host-side recipe of the sequence generator: RNG streams,
// row-mixing plan, hashes, subset layouts of the cyclic operator.
//
// These restate the reference driver's host code (cuda/cuprym_cyclic.cu) so
// that the files this library writes carry the same Dr/Dc/V/plan bits as the
// reference's for the same instance, seed and sector.

#include <algorithm>
#include <cstring>
#include <mutex>
#include <random>
#include <unordered_map>

#include "pgw_internal.h"

namespace pgw {

static thread_local std::string g_last_error = "no error";
void set_last_error(const std::string& msg) { g_last_error = msg; }

const uint64_t* host_binom_table() {
    static const std::vector<uint64_t> t = [] {
        std::vector<uint64_t> c(64 * 64, 0);
        for (int a = 0; a < 64; ++a) {
            c[a * 64] = 1;
            for (int b = 1; b <= a; ++b) c[a * 64 + b] = c[(a - 1) * 64 + b - 1] + (b < a ? c[(a - 1) * 64 + b] : 0);
        }
        return c;
    }();
    return t.data();
}

uint64_t host_binom(uint32_t n, uint32_t k) {
    if (k > n) return 0;
    if (k > n - k) k = n - k;
    unsigned __int128 out = 1;
    for (uint32_t i = 0; i < k; ++i) out = (out * (n - i)) / (i + 1);
    require(out <= (unsigned __int128)UINT64_MAX, "binomial coefficient does not fit in u64");
    return (uint64_t)out;
}

// ---------------------------------------------------------------------------
// RNG (cuda/cuprym_cyclic.cu:1143-1218)
// ---------------------------------------------------------------------------

static uint64_t splitmix64_next(uint64_t& state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static uint64_t bounded_draw(uint64_t& state, uint64_t bound) {
    uint64_t threshold = (0ULL - bound) % bound;
    for (;;) {
        uint64_t v = splitmix64_next(state);
        if (v >= threshold) return v % bound;
    }
}

constexpr uint64_t kFnvOffset = 14695981039346656037ULL;
constexpr uint64_t kFnvPrime = 1099511628211ULL;
static uint64_t fnv_u8(uint64_t h, uint8_t v) { return (h ^ v) * kFnvPrime; }
static uint64_t fnv_u64(uint64_t h, uint64_t v) {
    for (int b = 0; b < 8; ++b) h = fnv_u8(h, (uint8_t)(v >> (8 * b)));
    return h;
}
static uint64_t fnv_str(uint64_t h, const char* s) {
    for (; *s; ++s) h = fnv_u8(h, (uint8_t)*s);
    return h;
}

}  // namespace pgw

using namespace pgw;

template <typename F>
static int guarded(F&& f) {
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

extern "C" {

const char* pgw_last_error(void) { return g_last_error.c_str(); }

uint64_t pgw_seeded_sector_value(uint64_t seed, uint64_t sector) {
    uint64_t x = seed ^ (0x9E3779B97F4A7C15ULL + (sector << 6) + (sector >> 2));
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBULL;
    x ^= x >> 31;
    return x;
}

int pgw_reference_random_state(uint64_t seed, uint64_t sector, uint32_t prime, uint64_t rows,
                               uint64_t cols, uint32_t block, uint32_t* h_dr, uint32_t* h_dc,
                               uint32_t* h_v) {
    return guarded([&] {
        require(prime > 2, "prime must exceed 2");
        // Exactly the reference's generator: std::mt19937_64 and
        // std::uniform_int_distribution<uint64_t> (cuda/matrices.h:44-56).
        std::mt19937_64 gen(pgw_seeded_sector_value(seed, sector));
        std::uniform_int_distribution<uint64_t> nonzero(1, (uint64_t)prime - 1);
        std::uniform_int_distribution<uint64_t> any(0, (uint64_t)prime - 1);
        for (uint64_t i = 0; i < rows; ++i) {
            uint32_t v = (uint32_t)nonzero(gen);
            if (h_dr) h_dr[i] = v;
        }
        for (uint64_t i = 0; i < cols; ++i) {
            uint32_t v = (uint32_t)nonzero(gen);
            if (h_dc) h_dc[i] = v;
        }
        for (uint64_t i = 0; i < cols * block; ++i) {
            uint32_t v = (uint32_t)any(gen);
            if (h_v) h_v[i] = v;
        }
    });
}

uint64_t pgw_rowmix_realized_seed(uint64_t seed, const char* op_name, uint64_t sector,
                                  uint64_t rows, uint32_t prime) {
    uint64_t h = kFnvOffset;
    h = fnv_str(h, "paired-sl2-target-v1");
    h = fnv_str(h, op_name ? op_name : "");
    h = fnv_u64(h, seed);
    h = fnv_u64(h, sector);
    h = fnv_u64(h, rows);
    h = fnv_u64(h, prime);
    return pgw_seeded_sector_value(h, sector);
}

int pgw_rowmix_plan(uint64_t realized_seed, uint64_t rows, uint32_t rounds, uint32_t prime,
                    pgw_rowmix_pair* h_pairs) {
    return guarded([&] {
        require(rows <= UINT32_MAX, "rowmix target coordinate does not fit uint32");
        require(prime > 2, "prime must exceed 2");
        std::vector<uint32_t> perm(rows);
        for (uint64_t i = 0; i < rows; ++i) perm[i] = (uint32_t)i;
        uint64_t state = realized_seed;
        uint64_t pc = rows / 2;
        for (uint32_t round = 0; round < rounds; ++round) {
            // descending Fisher-Yates, then consecutive pairs with nonzero a, b
            for (uint64_t i = rows; i > 1; --i) {
                uint64_t j = bounded_draw(state, i);
                std::swap(perm[i - 1], perm[j]);
            }
            pgw_rowmix_pair* out = h_pairs + (uint64_t)round * pc;
            for (uint64_t q = 0; q < pc; ++q) {
                out[q].first = perm[2 * q];
                out[q].second = perm[2 * q + 1];
                out[q].a = (uint32_t)(1 + bounded_draw(state, prime - 1));
                out[q].b = (uint32_t)(1 + bounded_draw(state, prime - 1));
            }
        }
    });
}

uint64_t pgw_rowmix_plan_hash(uint64_t rows, uint32_t rounds, uint64_t seed,
                              uint64_t realized_seed, const pgw_rowmix_pair* h_pairs) {
    uint64_t h = kFnvOffset;
    h = fnv_str(h, "paired-sl2-target-v1");
    h = fnv_u64(h, rows);
    h = fnv_u64(h, rounds);
    h = fnv_u64(h, seed);
    h = fnv_u64(h, realized_seed);
    uint64_t pc = rows / 2;
    for (uint32_t round = 0; round < rounds; ++round) {
        h = fnv_u64(h, pc);
        for (uint64_t q = 0; q < pc; ++q) {
            const pgw_rowmix_pair& pr = h_pairs[(uint64_t)round * pc + q];
            h = fnv_u64(h, pr.first);
            h = fnv_u64(h, pr.second);
            h = fnv_u64(h, pr.a);
            h = fnv_u64(h, pr.b);
        }
    }
    return h;
}

uint64_t pgw_fnv1a64_u32(const uint32_t* h_values, uint64_t count) {
    uint64_t h = fnv_u64(kFnvOffset, count);
    for (uint64_t i = 0; i < count; ++i) h = fnv_u64(h, h_values[i]);
    return h;
}

int pgw_synthetic_table(uint64_t seed, uint32_t prime, uint64_t count, uint32_t* h_out) {
    return guarded([&] {
        require(prime > 2, "prime must exceed 2");
        uint64_t state = seed;
        for (uint64_t i = 0; i < count; ++i) h_out[i] = (uint32_t)bounded_draw(state, prime);
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// cyclic operator layouts (cuda/cuprym_cyclic.cu:611-819)
// ---------------------------------------------------------------------------

namespace pgw {

void build_char_map(const std::vector<uint16_t>& w, uint32_t r, std::vector<uint32_t>& off,
                    std::vector<uint16_t>& idx) {
    off.assign(r + 1, 0);
    idx.clear();
    for (uint32_t ch = 0; ch < r; ++ch) {
        off[ch] = (uint32_t)idx.size();
        for (size_t j = 0; j < w.size(); ++j)
            if (w[j] == ch) idx.push_back((uint16_t)j);
    }
    off[r] = (uint32_t)idx.size();
}

static uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t r) { return (a + r - (b % r)) % r; }

void build_phi_host(PhiHost& h, uint32_t nI, uint32_t nB, uint32_t nA, uint32_t r,
                    const uint32_t* mu, const uint16_t* wV, const uint16_t* wA0,
                    const uint16_t* wA1, const std::vector<uint32_t>& original, uint32_t degree,
                    uint32_t sector) {
    const uint32_t ln = (uint32_t)original.size();
    require(ln < 64, "implicit operator supports local_n < 64");
    require(degree >= 1 && degree <= ln, "source exterior degree is invalid");
    h.nI = nI; h.nB = nB; h.nA = nA; h.r = r;
    h.mu.assign(mu, mu + (size_t)nI * nB * nA);
    h.wV.assign(wV, wV + nI);
    h.wA0.assign(wA0, wA0 + nA);
    h.wA1.assign(wA1, wA1 + nB);
    build_char_map(h.wA0, r, h.a0_off, h.a0_idx);
    build_char_map(h.wA1, r, h.a1_off, h.a1_idx);
    HostBinom C;
    h.row_subsets = host_binom(ln, degree - 1);
    h.col_subsets = host_binom(ln, degree);
    require(h.row_subsets < (1ULL << 32) && h.col_subsets < (1ULL << 32),
            "subset count does not fit uint32");
    h.row_width = ln - (degree - 1);
    h.col_width = degree;
    auto weight_of = [&](uint64_t mask) {
        uint32_t s = 0;
        for (uint32_t j = 0; j < ln; ++j)
            if (mask >> j & 1ULL) s += wV[original[j]];
        return (uint16_t)(s % r);
    };
    h.row_subset_w.resize(h.row_subsets);
    h.col_subset_w.resize(h.col_subsets);
    h.row_inc_code.resize(h.row_subsets * h.row_width);
    h.row_inc_rank.resize(h.row_subsets * h.row_width);
    h.col_inc_code.resize(h.col_subsets * h.col_width);
    h.col_inc_rank.resize(h.col_subsets * h.col_width);
    // each subset's entries are its own: the subsets split across host threads
    parallel_rows(h.row_subsets, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t s = lo; s < hi; ++s) {
            uint64_t mask = unrank_mask(s, (int)ln, (int)degree - 1, C);
            h.row_subset_w[s] = weight_of(mask);
            uint32_t q = 0;
            for (uint32_t li = 0; li < ln; ++li) {
                if (mask >> li & 1ULL) continue;
                uint32_t position = (uint32_t)__builtin_popcountll(mask & ((1ULL << li) - 1));
                h.row_inc_code[s * h.row_width + q] = original[li] | ((position & 1u) << 31);
                h.row_inc_rank[s * h.row_width + q] =
                    (uint32_t)rank_mask(mask | (1ULL << li), (int)ln, (int)degree, C);
                ++q;
            }
        }
    });
    parallel_rows(h.col_subsets, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t s = lo; s < hi; ++s) {
            uint64_t mask = unrank_mask(s, (int)ln, (int)degree, C);
            h.col_subset_w[s] = weight_of(mask);
            uint32_t q = 0;
            for (uint32_t li = 0; li < ln; ++li) {
                if (!(mask >> li & 1ULL)) continue;
                h.col_inc_code[s * h.col_width + q] = original[li] | ((q & 1u) << 31);
                h.col_inc_rank[s * h.col_width + q] =
                    (uint32_t)rank_mask(mask & ~(1ULL << li), (int)ln, (int)degree - 1, C);
                ++q;
            }
        }
    });
    // sector layout
    h.row_off.assign(h.row_subsets + 1, 0);
    h.col_off.assign(h.col_subsets + 1, 0);
    h.row_rank.clear(); h.row_beta.clear(); h.col_rank.clear(); h.col_a.clear();
    for (uint64_t s = 0; s < h.row_subsets; ++s) {
        h.row_off[s] = (uint32_t)h.row_rank.size();
        uint32_t ch = sub_mod(sector, h.row_subset_w[s], r);
        for (uint32_t idx = h.a1_off[ch]; idx < h.a1_off[ch + 1]; ++idx) {
            h.row_rank.push_back((uint32_t)s);
            h.row_beta.push_back(h.a1_idx[idx]);
        }
    }
    h.row_off[h.row_subsets] = (uint32_t)h.row_rank.size();
    for (uint64_t s = 0; s < h.col_subsets; ++s) {
        h.col_off[s] = (uint32_t)h.col_rank.size();
        uint32_t ch = sub_mod(sector, h.col_subset_w[s], r);
        for (uint32_t idx = h.a0_off[ch]; idx < h.a0_off[ch + 1]; ++idx) {
            h.col_rank.push_back((uint32_t)s);
            h.col_a.push_back(h.a0_idx[idx]);
        }
    }
    h.col_off[h.col_subsets] = (uint32_t)h.col_rank.size();
}

void build_cyclic_layout(CyclicLayout& L, const pgw_operator_desc& d) {
    require(d.genus >= 6 && d.genus % 2 == 0, "cyclic operator expects an even genus >= 6");
    require(d.order >= 1, "cyclic order must be positive");
    require(d.sector < d.order, "sector must be in 0..r-1");
    require(d.mu && d.weight_v && d.weight_a0 && d.weight_a1 && d.kernel_weights &&
                d.right_inverse && d.kernel_basis,
            "cyclic operator description is missing arrays");
    const uint32_t g = d.genus, n = g - 3, m = g / 2, r = d.order;
    require(d.w_index < n, "elimination w_V_index is outside V");
    require(d.kernel_dim >= 1, "elimination kernel basis is empty");
    for (uint32_t i = 0; i < n; ++i) require(d.weight_v[i] < r, "V weight outside 0..r-1");
    for (uint32_t i = 0; i < g; ++i) require(d.weight_a0[i] < r, "A0 weight outside 0..r-1");
    for (uint32_t i = 0; i < n; ++i) require(d.weight_a1[i] < r, "A1 weight outside 0..r-1");
    for (uint32_t i = 0; i < d.kernel_dim; ++i)
        require(d.kernel_weights[i] < r, "kernel weight outside 0..r-1");
    std::vector<uint32_t> u;
    for (uint32_t i = 0; i < n; ++i)
        if (i != d.w_index) u.push_back(i);
    const uint32_t y_sector = sub_mod(d.sector, d.w_weight, r);
    build_phi_host(L.dm, n, n, g, r, d.mu, d.weight_v, d.weight_a0, d.weight_a1, u, m, d.sector);
    build_phi_host(L.d2, n, n, g, r, d.mu, d.weight_v, d.weight_a0, d.weight_a1, u, m - 1, y_sector);
    std::vector<uint16_t> kw(d.kernel_weights, d.kernel_weights + d.kernel_dim);
    std::vector<uint32_t> k_off;
    std::vector<uint16_t> k_idx;
    build_char_map(kw, r, k_off, k_idx);
    const uint64_t subsets = L.dm.row_subsets;  // (m-1)-subsets of U
    L.z_off.assign(subsets + 1, 0);
    L.z_rank.clear();
    L.z_kappa.clear();
    for (uint64_t s = 0; s < subsets; ++s) {
        L.z_off[s] = (uint32_t)L.z_rank.size();
        uint32_t ch = sub_mod(y_sector, L.dm.row_subset_w[s], r);
        for (uint32_t idx = k_off[ch]; idx < k_off[ch + 1]; ++idx) {
            L.z_rank.push_back((uint32_t)s);
            L.z_kappa.push_back(k_idx[idx]);
        }
    }
    L.z_off[subsets] = (uint32_t)L.z_rank.size();
    // validate_elimination_data (cuda/cuprym_cyclic.cu:423-481): mu_w R = I, mu_w B = 0
    const uint64_t p = d.prime;
    for (uint32_t beta = 0; beta < n; ++beta) {
        for (uint32_t gamma = 0; gamma < n; ++gamma) {
            uint64_t acc = 0;
            for (uint32_t a = 0; a < g; ++a)
                acc = (acc + (uint64_t)d.mu[((uint64_t)d.w_index * n + beta) * g + a] *
                                 d.right_inverse[(uint64_t)a * n + gamma]) % p;
            require(acc == (beta == gamma ? 1u : 0u), "elimination right inverse check mu_w*R=I failed");
        }
        for (uint32_t kk = 0; kk < d.kernel_dim; ++kk) {
            uint64_t acc = 0;
            for (uint32_t a = 0; a < g; ++a)
                acc = (acc + (uint64_t)d.mu[((uint64_t)d.w_index * n + beta) * g + a] *
                                 d.kernel_basis[(uint64_t)a * d.kernel_dim + kk]) % p;
            require(acc == 0, "elimination kernel check mu_w*B=0 failed");
        }
    }
    L.x_cols = L.dm.col_rank.size();
    L.z_count = L.z_rank.size();
    L.cols = L.x_cols + L.z_count;
    L.rows = L.d2.row_rank.size();
}

}  // namespace pgw

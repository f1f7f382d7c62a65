// pgw_cpu.cpp -- the exact host apply of the cyclic operators.
//
// S X = Dc F^T P^T Dr P F Dc X on a host block, evaluated matrix-free from the
// sector layout step by step in the order of the reference's CPU apply
// (cpu_apply_eliminated_square, cuda/cuprym_cyclic.cu:1591-1966): scale by Dc;
// t = D_m x; y = B z - R t; target = D_(m-1) y; the row-mixing rounds, Dr and
// the transposed rounds; then the adjoint chain D_(m-1)^T, R^T / B^T, D_m^T
// and Dc.  It does not go through the composed sparse maps, their transposes
// or the device, so comparing the device apply with it checks the map
// composition, the SELL form and the kernels.  Deformation variants follow
// the same chain with their dense columns (-dense for the augmented variant,
// +dense for the replacement, whose two dropped base columns read as zero).
// Every product is reduced at once: the result is the canonical residue.

#include <algorithm>

#include "pgw_internal.h"

namespace pgw {

namespace {

inline uint32_t mulm(uint32_t a, uint32_t b, uint32_t p) { return (uint32_t)((uint64_t)a * b % p); }
inline uint32_t addm(uint32_t a, uint32_t b, uint32_t p) {
    const uint32_t s = a + b;  // a, b < p < 2^31
    return s >= p ? s - p : s;
}
// acc[j] += c * x[j] for j < w
inline void axpy(uint32_t* acc, uint32_t c, const uint32_t* x, uint32_t w, uint32_t p) {
    for (uint32_t j = 0; j < w; ++j) acc[j] = addm(acc[j], mulm(c, x[j], p), p);
}

// one row-mixing round on (first, second): forward [[1,a],[b,1+ab]],
// transpose [[1,b],[a,1+ab]] (cpu_apply_rowmix_forward / _transpose)
void rowmix_round(uint32_t* t, const pgw_rowmix_pair* pr, uint64_t pc, uint32_t w, uint32_t p,
                  bool transpose) {
    for (uint64_t q = 0; q < pc; ++q) {
        const uint32_t a = pr[q].a, b = pr[q].b;
        const uint32_t opab = addm(1, mulm(a, b, p), p);
        const uint32_t c01 = transpose ? b : a, c10 = transpose ? a : b;
        uint32_t* f = t + (uint64_t)pr[q].first * w;
        uint32_t* s = t + (uint64_t)pr[q].second * w;
        for (uint32_t j = 0; j < w; ++j) {
            const uint32_t xf = f[j], xs = s[j];
            f[j] = addm(xf, mulm(c01, xs, p), p);
            s[j] = addm(mulm(c10, xf, p), mulm(opab, xs, p), p);
        }
    }
}

}  // namespace

void cyclic_apply_host(const CyclicLayout& L, const pgw_operator_desc& d, uint64_t cols, const uint32_t* dr,
                       const uint32_t* dc, uint32_t rounds, const pgw_rowmix_pair* pairs, const uint32_t* in,
                       uint32_t w, uint32_t* out) {
    const uint32_t p = d.prime;
    const PhiHost& dm = L.dm;
    const PhiHost& d2 = L.d2;
    const uint32_t nv = d.genus - 3, kdim = d.kernel_dim;
    const uint64_t x_cols = L.x_cols, base_cols = L.cols;
    const uint64_t t_rows = dm.row_rank.size(), y_rows = d2.col_rank.size(), rows = d2.row_rank.size();
    const bool aug = d.kind == PGW_KIND_CYCLIC_AUGMENTED;
    const bool rep = d.kind == PGW_KIND_CYCLIC_REPLACEMENT;
    const uint32_t dcols = (aug || rep) ? d.dense_cols : 0;
    uint32_t drop0 = 0, drop1 = 0;
    if (rep) {
        drop0 = std::min(d.drop_cols[0], d.drop_cols[1]);
        drop1 = std::max(d.drop_cols[0], d.drop_cols[1]);
    }
    // source column of base position f, ~0 for a column the replacement drops
    auto base_to_x = [&](uint64_t f) -> uint64_t {
        if (!rep) return f;
        if (f == drop0 || f == drop1) return ~0ULL;
        return f - (f > drop0 ? 1 : 0) - (f > drop1 ? 1 : 0);
    };
    const uint64_t dense_x0 = aug ? base_cols : base_cols - 2;  // first dense source column

    // Dc X, then the base view of it (dropped columns read as zero)
    std::vector<uint32_t> xs(cols * w);
    for (uint64_t c = 0; c < cols; ++c)
        for (uint32_t j = 0; j < w; ++j) xs[c * w + j] = mulm(dc[c], in[c * w + j], p);
    std::vector<uint32_t> zero(w, 0);
    auto base_row = [&](uint64_t f) -> const uint32_t* {
        const uint64_t x = base_to_x(f);
        return x == ~0ULL ? zero.data() : xs.data() + x * w;
    };

    // t = D_m x (the x part of the base columns)
    std::vector<uint32_t> t(t_rows * w, 0);
    for (uint64_t r = 0; r < t_rows; ++r)
        phi_forward_row(dm, r, p, [&](uint64_t f, uint32_t c) { axpy(&t[r * w], c, base_row(f), w, p); });

    // y = B z - R t: the z entries of y's subset (kernel basis) and the t rows
    // of that subset (right inverse), cyclic_elim_z_to_y / _subtract_Rt
    std::vector<uint32_t> y(y_rows * w, 0);
    for (uint64_t r = 0; r < y_rows; ++r) {
        const uint32_t s = d2.col_rank[r], a = d2.col_a[r];
        for (uint32_t q = L.z_off[s]; q < L.z_off[s + 1]; ++q)
            axpy(&y[r * w], d.kernel_basis[(uint64_t)a * kdim + L.z_kappa[q]], base_row(x_cols + q), w, p);
        for (uint32_t tr = dm.row_off[s]; tr < dm.row_off[s + 1]; ++tr) {
            const uint32_t rv = d.right_inverse[(uint64_t)a * nv + dm.row_beta[tr]];
            if (rv) axpy(&y[r * w], p - rv, &t[(uint64_t)tr * w], w, p);
        }
    }

    // target = D_(m-1) y, with the dense columns of a deformation variant
    std::vector<uint32_t> tg(rows * w, 0);
    for (uint64_t r = 0; r < rows; ++r) {
        phi_forward_row(d2, r, p, [&](uint64_t yc, uint32_t c) { axpy(&tg[r * w], c, &y[yc * w], w, p); });
        for (uint32_t j = 0; j < dcols; ++j) {
            const uint32_t v = d.dense_block[r * dcols + j];
            if (v) axpy(&tg[r * w], aug ? p - v : v, &xs[(dense_x0 + j) * w], w, p);
        }
    }

    // P, Dr, P^T
    const uint64_t pc = rows / 2;
    if (pc)
        for (uint32_t rd = 0; rd < rounds; ++rd) rowmix_round(tg.data(), pairs + (uint64_t)rd * pc, pc, w, p, false);
    for (uint64_t r = 0; r < rows; ++r)
        for (uint32_t j = 0; j < w; ++j) tg[r * w + j] = mulm(dr[r], tg[r * w + j], p);
    if (pc)
        for (uint32_t rd = rounds; rd-- > 0;) rowmix_round(tg.data(), pairs + (uint64_t)rd * pc, pc, w, p, true);

    // adjoint chain: h = D_(m-1)^T target (and the dense columns' share)
    std::vector<uint32_t> h(y_rows * w, 0), hd((uint64_t)dcols * w, 0);
    for (uint64_t r = 0; r < rows; ++r) {
        phi_forward_row(d2, r, p, [&](uint64_t yc, uint32_t c) { axpy(&h[yc * w], c, &tg[r * w], w, p); });
        for (uint32_t j = 0; j < dcols; ++j) {
            const uint32_t v = d.dense_block[r * dcols + j];
            if (v) axpy(&hd[(uint64_t)j * w], aug ? p - v : v, &tg[r * w], w, p);
        }
    }
    // R^T h into the t rows, B^T h into the z positions
    std::vector<uint32_t> tadj(t_rows * w, 0), zadj(L.z_count * w, 0);
    for (uint64_t r = 0; r < y_rows; ++r) {
        const uint32_t s = d2.col_rank[r], a = d2.col_a[r];
        for (uint32_t q = L.z_off[s]; q < L.z_off[s + 1]; ++q)
            axpy(&zadj[(uint64_t)q * w], d.kernel_basis[(uint64_t)a * kdim + L.z_kappa[q]], &h[r * w], w, p);
        for (uint32_t tr = dm.row_off[s]; tr < dm.row_off[s + 1]; ++tr) {
            const uint32_t rv = d.right_inverse[(uint64_t)a * nv + dm.row_beta[tr]];
            if (rv) axpy(&tadj[(uint64_t)tr * w], p - rv, &h[r * w], w, p);
        }
    }
    // D_m^T into the x positions
    std::vector<uint32_t> xadj(x_cols * w, 0);
    for (uint64_t r = 0; r < t_rows; ++r)
        phi_forward_row(dm, r, p, [&](uint64_t f, uint32_t c) { axpy(&xadj[f * w], c, &tadj[r * w], w, p); });

    // assemble in source-column order and scale by Dc
    std::fill(out, out + cols * w, 0u);
    auto put = [&](uint64_t x, const uint32_t* v) {
        for (uint32_t j = 0; j < w; ++j) out[x * w + j] = mulm(dc[x], v[j], p);
    };
    for (uint64_t f = 0; f < x_cols; ++f) {
        const uint64_t x = base_to_x(f);
        if (x != ~0ULL) put(x, &xadj[f * w]);
    }
    for (uint64_t q = 0; q < L.z_count; ++q) {
        const uint64_t x = base_to_x(x_cols + q);
        if (x != ~0ULL) put(x, &zadj[q * w]);
    }
    for (uint32_t j = 0; j < dcols; ++j) put(dense_x0 + j, &hd[(uint64_t)j * w]);
}

}  // namespace pgw

using namespace pgw;

extern "C" {

int pgw_cyclic_apply_cpu(const pgw_operator_desc* desc, const uint32_t* h_row_precond,
                         const uint32_t* h_col_precond, uint32_t rowmix_rounds, const pgw_rowmix_pair* h_pairs,
                         const uint32_t* h_in, uint32_t width, uint32_t* h_out) {
    try {
        validate_desc_common(desc);
        require(desc->kind == PGW_KIND_CYCLIC_ELIMINATED || is_deformation(desc->kind),
                "the host apply covers the cyclic operators");
        require(h_in && h_out, "null input or output block");
        require(width >= 1 && width <= 16, "block width must be in 1..16");
        CyclicLayout L;
        build_cyclic_layout(L, *desc);
        uint64_t cols = L.cols;
        if (is_deformation(desc->kind)) {
            uint32_t d0, d1;
            cols = deformation_cols(desc, L.cols, d0, d1);
            uint32_t lo, hi;
            min_max_u32(desc->dense_block, L.rows * desc->dense_cols, lo, hi);
            require(L.rows == 0 || hi < desc->prime, "dense block entries must be reduced modulo p");
        }
        const uint32_t g = desc->genus, nv = g - 3;
        for (uint64_t q = 0; q < (uint64_t)nv * nv * g; ++q)
            require(desc->mu[q] < desc->prime, "mu.values entry is not reduced modulo the instance modulus");
        validate_preconditioners(h_row_precond, L.rows, h_col_precond, cols, desc->prime);
        validate_rowmix_plan(L.rows, rowmix_rounds, h_pairs, desc->prime);
        uint32_t lo, hi;
        min_max_u32(h_in, cols * width, lo, hi);
        require(cols == 0 || hi < desc->prime, "input entries must be reduced modulo p");
        cyclic_apply_host(L, *desc, cols, h_row_precond, h_col_precond, rowmix_rounds, h_pairs, h_in, width,
                          h_out);
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

int pgw_deterministic_block(uint64_t rows, uint64_t cols, uint32_t prime, uint64_t salt, uint32_t* h_out) {
    if (!h_out || prime < 2) {
        set_last_error("null output or modulus below 2");
        return PGW_ERR_INVALID;
    }
    uint64_t state = salt ^ (0xD6E8FEB86659FD93ULL + rows * 1315423911ULL + cols * 2654435761ULL);
    for (uint64_t q = 0; q < rows * cols; ++q) {
        uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        h_out[q] = (uint32_t)((z ^ (z >> 31)) % prime);
    }
    return PGW_OK;
}

}  // extern "C"

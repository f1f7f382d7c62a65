// pgw_sparse.cpp -- the eliminated cyclic operator (and its deformation
// variants) as a short chain of explicit sparse maps.
//
// The reference applies F = [ -D_(m-1) R D_m | D_(m-1) B ] matrix-free: every
// output entry walks its subset's incidences, character blocks and the
// elimination matrices through chains of dependent lookups
// (cuda/cyclic_cuda_helpers.h:145-266, 718-895; the CPU twin
// cuda/cuprym_cyclic.cu:1591-1966).  Here the host expands those same index
// rules once per operator into explicit sparse rows, with the column
// preconditioner, the elimination and the row mixing folded in, so that one
// block-apply S = Dc F^T P^T Dr P F Dc is five sparse products:
//   M1: x -> y'   (Dc, D_m, -R and B composed; y' = y plus the dense-column inputs)
//   M2: y' -> t   (D_(m-1), plus the dense columns of a deformation variant)
//   M3: t -> t    (P^T Dr P of the row-mixing rounds)
//   M2^T, M1^T    (transposes, so that S is exactly G^T Dr G with G = P F Dc).
// All coefficients are canonical residues computed exactly on the host; the
// device sums are exact, so the result is the same residue as the
// matrix-free apply's.

#include <algorithm>
#include <thread>

#include "pgw_internal.h"

namespace pgw {

namespace {

inline uint32_t mulm(uint64_t a, uint64_t b, uint64_t p) { return (uint32_t)(a * b % p); }


// Rows built independently (row r's entries into a per-row list), then
// flattened into CSR with duplicate columns summed mod p and zeros dropped.
struct RowLists {
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> rows;
    explicit RowLists(uint64_t n) : rows(n) {}
};

CsrHost flatten(RowLists& L, uint64_t cols, uint32_t p) {
    CsrHost M;
    M.rows = L.rows.size();
    M.cols = cols;
    // merge within rows (in parallel), then count
    parallel_rows(M.rows, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t r = lo; r < hi; ++r) {
            auto& v = L.rows[r];
            std::sort(v.begin(), v.end());
            size_t w = 0;
            for (size_t q = 0; q < v.size();) {
                const uint32_t c = v[q].first;
                uint64_t s = 0;
                for (; q < v.size() && v[q].first == c; ++q) s += v[q].second;  // < 2^32 terms of < 2^31
                const uint32_t val = (uint32_t)(s % p);
                if (val) v[w++] = {c, val};
            }
            v.resize(w);
        }
    });
    M.ptr.assign(M.rows + 1, 0);
    for (uint64_t r = 0; r < M.rows; ++r) M.ptr[r + 1] = M.ptr[r] + L.rows[r].size();
    M.col.resize(M.ptr[M.rows]);
    M.val.resize(M.ptr[M.rows]);
    parallel_rows(M.rows, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t r = lo; r < hi; ++r) {
            uint64_t k = M.ptr[r];
            for (const auto& e : L.rows[r]) {
                M.col[k] = e.first;
                M.val[k] = e.second;
                ++k;
            }
            std::vector<std::pair<uint32_t, uint32_t>>().swap(L.rows[r]);
        }
    });
    return M;
}

}  // namespace

CsrHost csr_transpose(const CsrHost& A) {
    CsrHost T;
    T.rows = A.cols;
    T.cols = A.rows;
    T.ptr.assign(T.rows + 1, 0);
    for (uint32_t c : A.col) ++T.ptr[(uint64_t)c + 1];
    for (uint64_t r = 0; r < T.rows; ++r) T.ptr[r + 1] += T.ptr[r];
    T.col.resize(A.col.size());
    T.val.resize(A.val.size());
    std::vector<uint64_t> at(T.ptr.begin(), T.ptr.end() - 1);
    for (uint64_t r = 0; r < A.rows; ++r)  // ascending source rows: sorted columns in T
        for (uint64_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k) {
            const uint64_t d = at[A.col[k]]++;
            T.col[d] = (uint32_t)r;
            T.val[d] = A.val[k];
        }
    return T;
}

SellHost csr_to_sell(const CsrHost& A) {
    SellHost S;
    S.rows = A.rows;
    S.cols = A.cols;
    const uint64_t nslices = (A.rows + kSellHeight - 1) / kSellHeight;
    S.slice.assign(nslices + 1, 0);
    for (uint64_t s = 0; s < nslices; ++s) {
        uint64_t w = 0;
        for (uint64_t r = s * kSellHeight; r < std::min(A.rows, (s + 1) * kSellHeight); ++r)
            w = std::max(w, A.ptr[r + 1] - A.ptr[r]);
        S.slice[s + 1] = S.slice[s] + w * kSellHeight;
    }
    S.nnz = A.ptr[A.rows];
    S.ent.assign(2 * S.slice[nslices], 0);  // padding: column 0, coefficient 0
    parallel_rows(nslices, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t s = lo; s < hi; ++s)
            for (uint64_t r = s * kSellHeight; r < std::min(A.rows, (s + 1) * kSellHeight); ++r) {
                const uint64_t lane = r - s * kSellHeight;
                for (uint64_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k) {
                    const uint64_t e = S.slice[s] + (k - A.ptr[r]) * kSellHeight + lane;
                    S.ent[2 * e] = A.col[k];
                    S.ent[2 * e + 1] = A.val[k];
                }
            }
    });
    return S;
}

void build_cyclic_maps(CyclicMaps& out, const CyclicLayout& L, const pgw_operator_desc& d,
                       uint64_t cols, const uint32_t* dc, const uint32_t* dr, uint32_t rounds,
                       const pgw_rowmix_pair* pairs) {
    const uint32_t p = d.prime;
    const PhiHost& dm = L.dm;
    const PhiHost& d2 = L.d2;
    const uint32_t nv = d.genus - 3, kdim = d.kernel_dim;
    const uint64_t base_cols = L.cols, x_cols = L.x_cols, y_rows = d2.col_rank.size();
    const bool aug = d.kind == PGW_KIND_CYCLIC_AUGMENTED;
    const bool rep = d.kind == PGW_KIND_CYCLIC_REPLACEMENT;
    const uint32_t dcols = (aug || rep) ? d.dense_cols : 0;
    uint32_t drop0 = 0, drop1 = 0;
    if (rep) {
        drop0 = std::min(d.drop_cols[0], d.drop_cols[1]);
        drop1 = std::max(d.drop_cols[0], d.drop_cols[1]);
    }
    // source column of base position f (~0: a dropped column of the replacement)
    auto base_to_x = [&](uint64_t f) -> uint64_t {
        if (!rep) return f;
        if (f == drop0 || f == drop1) return ~0ULL;
        return f - (f > drop0 ? 1 : 0) - (f > drop1 ? 1 : 0);
    };
    // source column of dense input j: after the base (augmented) or kept (replacement) columns
    const uint64_t dense_x0 = aug ? base_cols : base_cols - 2;
    require(!dcols || dense_x0 + dcols == cols, "deformation column count mismatch");
    require(dcols || base_cols == cols, "eliminated column count mismatch");

    // M1 rows: y (D_(m-1) columns) then one pass-through row per dense input
    RowLists m1(y_rows + dcols);
    parallel_rows(y_rows, [&](uint64_t lo, uint64_t hi) {
        std::vector<std::pair<uint32_t, uint32_t>> trow;
        for (uint64_t row = lo; row < hi; ++row) {
            auto& out_row = m1.rows[row];
            const uint32_t s = d2.col_rank[row], a = d2.col_a[row];
            // y = B z: z index q reads scaled input x_cols + q (cyclic_elim_z_to_y)
            for (uint32_t q = L.z_off[s]; q < L.z_off[s + 1]; ++q) {
                const uint64_t x = base_to_x(x_cols + q);
                if (x == ~0ULL) continue;
                const uint32_t coef = mulm(d.kernel_basis[(uint64_t)a * kdim + L.z_kappa[q]], dc[x], p);
                if (coef) out_row.emplace_back((uint32_t)x, coef);
            }
            // y -= R t: t rows of subset s (cyclic_elim_subtract_Rt), t = D_m (Dc x)
            for (uint32_t t = dm.row_off[s]; t < dm.row_off[s + 1]; ++t) {
                const uint32_t rv = d.right_inverse[(uint64_t)a * nv + dm.row_beta[t]];
                if (rv == 0) continue;
                const uint64_t neg_r = p - rv;
                phi_forward_row(dm, t, p, [&](uint64_t f, uint32_t coef) {
                    const uint64_t x = base_to_x(f);
                    if (x == ~0ULL) return;
                    const uint32_t c = mulm(mulm(neg_r, coef, p), dc[x], p);
                    if (c) out_row.emplace_back((uint32_t)x, c);
                });
            }
        }
    });
    for (uint32_t j = 0; j < dcols; ++j)
        m1.rows[y_rows + j].emplace_back((uint32_t)(dense_x0 + j), dc[dense_x0 + j]);
    out.m1 = flatten(m1, cols, p);

    // M2 rows: D_(m-1) y, plus -dense (augmented) or +dense (replacement)
    const uint64_t rows = d2.row_rank.size();
    RowLists m2(rows);
    parallel_rows(rows, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t row = lo; row < hi; ++row) {
            auto& out_row = m2.rows[row];
            phi_forward_row(d2, row, p, [&](uint64_t yc, uint32_t coef) {
                out_row.emplace_back((uint32_t)yc, coef);
            });
            for (uint32_t j = 0; j < dcols; ++j) {
                const uint32_t v = d.dense_block[row * dcols + j];
                if (v) out_row.emplace_back((uint32_t)(y_rows + j), aug ? p - v : v);
            }
        }
    });
    out.m2 = flatten(m2, y_rows + dcols, p);

    // M3 = P^T Dr P, P = P_(R-1) ... P_0, round maps [[1,a],[b,1+ab]] on (first, second)
    // (cyclic_rowmix_pair_kernel, cuda/cyclic_cuda_helpers.h:109-143)
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> P(rows);
    for (uint64_t r = 0; r < rows; ++r) P[r] = {{(uint32_t)r, 1u}};
    const uint64_t pc = rows / 2;
    for (uint32_t rd = 0; rd < rounds && pc; ++rd) {
        parallel_rows(pc, [&](uint64_t lo, uint64_t hi) {
            std::vector<std::pair<uint32_t, uint32_t>> nf, ns;
            for (uint64_t q = lo; q < hi; ++q) {
                const pgw_rowmix_pair& pr = pairs[(uint64_t)rd * pc + q];
                const uint64_t a = pr.a, b = pr.b, opab = (1 + a * b % p) % p;
                auto& F = P[pr.first];
                auto& S = P[pr.second];
                nf.clear();
                ns.clear();
                for (const auto& e : F) {
                    nf.emplace_back(e.first, e.second);
                    ns.emplace_back(e.first, mulm(b, e.second, p));
                }
                for (const auto& e : S) {
                    nf.emplace_back(e.first, mulm(a, e.second, p));
                    ns.emplace_back(e.first, mulm(opab, e.second, p));
                }
                F = nf;
                S = ns;
            }
        });
    }
    // rows of P^T: which mixed rows u read source row r, with P[u][r]
    std::vector<uint64_t> tp(rows + 1, 0);
    for (uint64_t u = 0; u < rows; ++u)
        for (const auto& e : P[u]) ++tp[(uint64_t)e.first + 1];
    for (uint64_t r = 0; r < rows; ++r) tp[r + 1] += tp[r];
    std::vector<std::pair<uint32_t, uint32_t>> pt(tp[rows]);
    {
        std::vector<uint64_t> at(tp.begin(), tp.end() - 1);
        for (uint64_t u = 0; u < rows; ++u)
            for (const auto& e : P[u]) pt[at[e.first]++] = {(uint32_t)u, e.second};
    }
    RowLists m3(rows);
    parallel_rows(rows, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t r = lo; r < hi; ++r) {
            auto& out_row = m3.rows[r];
            for (uint64_t k = tp[r]; k < tp[r + 1]; ++k) {
                const uint32_t u = pt[k].first;
                const uint32_t w = mulm(pt[k].second, dr[u], p);
                for (const auto& e : P[u]) out_row.emplace_back(e.first, mulm(w, e.second, p));
            }
        }
    });
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>>().swap(P);
    out.m3 = flatten(m3, rows, p);
}

std::vector<uint32_t> mix_order(uint64_t rows, uint32_t rounds, const pgw_rowmix_pair* pairs) {
    const uint64_t pc = rows / 2;
    // partner of each row in each round (~0: unpaired in that round)
    std::vector<uint32_t> partner((uint64_t)rounds * rows, ~0u);
    for (uint32_t r = 0; r < rounds; ++r)
        for (uint64_t q = 0; q < pc; ++q) {
            const pgw_rowmix_pair& pr = pairs[(uint64_t)r * pc + q];
            partner[(uint64_t)r * rows + pr.first] = pr.second;
            partner[(uint64_t)r * rows + pr.second] = pr.first;
        }
    std::vector<uint32_t> order;
    order.reserve(rows);
    std::vector<char> seen(rows, 0);
    // from `start`, step to an unseen partner (first round first) until none
    // is left; with at most two rounds that traces a whole path or cycle
    auto walk = [&](uint64_t start) {
        uint64_t cur = start;
        for (;;) {
            seen[cur] = 1;
            order.push_back((uint32_t)cur);
            uint64_t next = ~0ULL;
            for (uint32_t r = 0; r < rounds && next == ~0ULL; ++r) {
                const uint32_t c = partner[(uint64_t)r * rows + cur];
                if (c != ~0u && !seen[c]) next = c;
            }
            if (next == ~0ULL) return;
            cur = next;
        }
    };
    // path ends (a row unpaired in some round) first, so paths are walked end to end
    for (uint64_t v = 0; v < rows; ++v) {
        if (seen[v]) continue;
        bool end = false;
        for (uint32_t r = 0; r < rounds; ++r) end |= partner[(uint64_t)r * rows + v] == ~0u;
        if (end) walk(v);
    }
    for (uint64_t v = 0; v < rows; ++v)
        if (!seen[v]) walk(v);
    return order;
}

MixChains mix_chains(uint64_t rows, uint32_t rounds, const pgw_rowmix_pair* pairs, const uint32_t* dr) {
    require(rounds >= 1 && rounds <= 2, "row-mixing chains need one or two rounds");
    require(rows < (1ULL << 31), "row-mixing chains need fewer than 2^31 rows");
    const uint64_t pc = rows / 2;
    // partner row in each round, or kNone; the high bit of a row's first
    // word marks it walked, so a step of the walk touches one cache line: the
    // next row's, which it reads anyway
    constexpr uint32_t kNone = 0x7fffffffu, kSeen = 0x80000000u;
    std::vector<uint32_t> nbr(2 * rows, kNone);
    for (uint32_t r = 0; r < rounds; ++r)
        parallel_rows(pc, [&](uint64_t lo, uint64_t hi) {
            for (uint64_t q = lo; q < hi; ++q) {
                const pgw_rowmix_pair& pr = pairs[(uint64_t)r * pc + q];
                nbr[2 * (uint64_t)pr.first + r] = pr.second;
                nbr[2 * (uint64_t)pr.second + r] = pr.first;
            }
        });
    auto partner = [&](uint64_t v, uint32_t r) { return nbr[2 * v + r] & ~kSeen; };
    auto seen = [&](uint64_t v) { return (nbr[2 * v] & kSeen) != 0; };
    // the walk: chain positions (row, round of the link to the next position,
    // ghost flag); the pair coefficients are filled in afterwards, in parallel
    std::vector<uint32_t> row;
    std::vector<uint8_t> link;
    row.reserve(rows + 64);
    link.reserve(rows + 64);
    std::vector<uint32_t> comp;
    std::vector<uint8_t> clink;
    auto walk = [&](uint64_t start) {  // returns the round closing a cycle, or kMixNone
        comp.clear();
        clink.clear();
        uint64_t cur = start;
        for (;;) {
            nbr[2 * cur] |= kSeen;
            comp.push_back((uint32_t)cur);
            uint64_t next = ~0ULL;
            uint32_t via = kMixNone;
            for (uint32_t r = 0; r < rounds && next == ~0ULL; ++r) {
                const uint32_t q = partner(cur, r);
                if (q != kNone && !seen(q)) {
                    next = q;
                    via = r;
                }
            }
            if (next == ~0ULL) break;
            clink.push_back((uint8_t)via);
            cur = next;
        }
        if (comp.size() >= 2)
            for (uint32_t r = 0; r < rounds; ++r)
                if (partner(comp.back(), r) == comp.front() && r != clink.back()) return r;
        return kMixNone;
    };
    auto emit_component = [&](uint32_t close) {
        const uint64_t L = comp.size();
        if (close == kMixNone) {  // a path
            for (uint64_t k = 0; k < L; ++k) {
                row.push_back(comp[k]);
                link.push_back((uint8_t)(k + 1 < L ? clink[k] : kMixNone));
            }
            return;
        }
        // a cycle: edge k joins positions k and k + 1 (mod L), edge L - 1 is the
        // closure; kMixHalo ghost positions on either side
        for (int64_t k = -(int64_t)kMixHalo; k < (int64_t)(L + kMixHalo); ++k) {
            const uint64_t m = (uint64_t)((k % (int64_t)L + (int64_t)L) % (int64_t)L);
            const bool last = k + 1 == (int64_t)(L + kMixHalo);
            const uint32_t rd = last ? kMixNone : (m + 1 < L ? clink[m] : close);
            row.push_back(comp[m]);
            link.push_back((uint8_t)(rd | ((k < 0 || k >= (int64_t)L) ? kMixGhost : 0u)));
        }
    };
    // paths from their ends first, then the cycles
    for (uint64_t v = 0; v < rows; ++v) {
        if (seen(v)) continue;
        bool end = false;
        for (uint32_t r = 0; r < rounds; ++r) end |= partner(v, r) == kNone;
        if (end) emit_component(walk(v));
    }
    for (uint64_t v = 0; v < rows; ++v)
        if (!seen(v)) emit_component(walk(v));
    std::vector<uint32_t>().swap(nbr);
    // pair index of each row in each round, for the coefficients
    std::vector<uint32_t> pidx(2 * rows, ~0u);
    for (uint32_t r = 0; r < rounds; ++r)
        parallel_rows(pc, [&](uint64_t lo, uint64_t hi) {
            for (uint64_t q = lo; q < hi; ++q) {
                const pgw_rowmix_pair& pr = pairs[(uint64_t)r * pc + q];
                pidx[2 * (uint64_t)pr.first + r] = (uint32_t)q;
                pidx[2 * (uint64_t)pr.second + r] = (uint32_t)q;
            }
        });
    MixChains c;
    const uint64_t n = row.size();
    c.a.assign(n, 0);
    c.b.assign(n, 0);
    c.dr.assign(n, 0);
    c.flag.assign(n, 0);
    parallel_rows(n, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) {
            const uint32_t v = row[i], rd = link[i] & 3u;
            uint8_t f = link[i];
            if (rd != kMixNone) {
                const pgw_rowmix_pair& pr = pairs[(uint64_t)rd * pc + pidx[2 * (uint64_t)v + rd]];
                require(i + 1 < n && (pr.first == row[i + 1] || pr.second == row[i + 1]),
                        "row-mixing chain link is not a pair");
                c.a[i] = pr.a;
                c.b[i] = pr.b;
                if (pr.first == v) f |= kMixFirst;
            }
            c.flag[i] = f;
            c.dr[i] = dr[v];
        }
    });
    c.row = std::move(row);
    return c;
}

CsrHost permute_square(const CsrHost& A, const std::vector<uint32_t>& order) {
    require(A.rows == A.cols && order.size() == A.rows, "permutation does not match the map");
    std::vector<uint32_t> pos(A.rows);
    for (uint64_t i = 0; i < A.rows; ++i) pos[order[i]] = (uint32_t)i;
    CsrHost P;
    P.rows = P.cols = A.rows;
    P.ptr.assign(A.rows + 1, 0);
    for (uint64_t i = 0; i < A.rows; ++i) P.ptr[i + 1] = P.ptr[i] + (A.ptr[order[i] + 1] - A.ptr[order[i]]);
    P.col.resize(P.ptr[A.rows]);
    P.val.resize(P.ptr[A.rows]);
    parallel_rows(A.rows, [&](uint64_t lo, uint64_t hi) {
        std::vector<std::pair<uint32_t, uint32_t>> row;
        for (uint64_t i = lo; i < hi; ++i) {
            row.clear();
            for (uint64_t k = A.ptr[order[i]]; k < A.ptr[order[i] + 1]; ++k) row.emplace_back(pos[A.col[k]], A.val[k]);
            std::sort(row.begin(), row.end());
            for (size_t t = 0; t < row.size(); ++t) {
                P.col[P.ptr[i] + t] = row[t].first;
                P.val[P.ptr[i] + t] = row[t].second;
            }
        }
    });
    return P;
}

}  // namespace pgw

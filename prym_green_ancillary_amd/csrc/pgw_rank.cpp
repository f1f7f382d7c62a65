// pgw_rank.cpp -- rank lower bound and matrix generator from a symmetric block
// Wiedemann sequence: the consumer of the WDM files (host code, threads).
//
// Semantics of the reference's BCW stage (third_party/bcw/src/rank_pipeline.rs:
// 119-307, sigma_basis.rs:37-110,233-324, main.rs:31-100):
//   * the initial V^T V term is dropped; term t >= 1 is the b x b symmetric
//     matrix S_t, filled from its packed upper triangle;
//   * G(x) = [ sum_t S_(t+1) x^t ; I_b ]  (2b x b), shift delta = (0^b, 1^b);
//   * a shifted order basis M with M G = 0 mod x^d, d = usable terms, built by
//     the divide-and-conquer PM-basis whose leaves are one elimination step:
//     rows stably sorted by delta, an LSP elimination of the degree-k residual
//     coefficient, pivot rows multiplied by x (delta + 1);
//   * rank lower bound = sum of the b smallest deltas, accepted when the gap
//     between the b-th and (b+1)-th smallest reaches the threshold (10).
// Every leaf performs the same row operations as the reference's `basis`
// (same stable order, same LSP pivot columns, same elimination factors), so
// the basis, and with it the generator file, is the reference's exactly; the
// split points of the recursion do not change the leaves.
//
// Polynomial matrix products run through three or four NTT primes and a
// Garner CRT back to p (schoolbook for short operands), so every prime
// 2 <= p < 2^31 works, p = 2^31 - 1 included (the reference's single-prime
// NTT admits only p^2 * 2^ceil(log2(2 max_nlen)) * b < 2305843009146585089,
// poly_mat_mul.rs:271-297).

#include <algorithm>
#include <atomic>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "pgw_op.h"

using namespace pgw;

namespace {

// ---------------------------------------------------------------------------
// parallel loop over [0, n) on up to `threads` threads (the caller included)
// ---------------------------------------------------------------------------
template <class F>
void parallel_for(uint64_t n, uint32_t threads, F&& f) {
    if (threads <= 1 || n <= 1) {
        for (uint64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<uint64_t> next{0};
    std::exception_ptr err;
    std::mutex err_mtx;
    auto body = [&] {
        try {
            for (uint64_t i; (i = next.fetch_add(1)) < n;) f(i);
        } catch (...) {
            std::lock_guard<std::mutex> g(err_mtx);
            if (!err) err = std::current_exception();
            next.store(n);
        }
    };
    const uint32_t t = (uint32_t)std::min<uint64_t>(threads, n);
    std::vector<std::thread> pool;
    pool.reserve(t - 1);
    for (uint32_t w = 1; w < t; ++w) pool.emplace_back(body);
    body();
    for (auto& th : pool) th.join();
    if (err) std::rethrow_exception(err);
}

uint32_t pow_mod(uint32_t a, uint64_t e, const ModP& m) {
    uint32_t r = 1 % m.p, x = a;
    while (e) {
        if (e & 1) r = mulmod(r, x, m);
        x = mulmod(x, x, m);
        e >>= 1;
    }
    return r;
}

uint32_t inv_mod(uint32_t a, const ModP& m) { return pow_mod(a, m.p - 2, m); }

// ---------------------------------------------------------------------------
// polynomial matrices: entry (i, j) holds `len` coefficients, ascending degree
// ---------------------------------------------------------------------------
struct PolyMat {
    uint32_t r = 0, c = 0;
    uint64_t len = 0;
    std::vector<uint32_t> a;

    PolyMat() = default;
    PolyMat(uint32_t r_, uint32_t c_, uint64_t len_)
        : r(r_), c(c_), len(len_), a((uint64_t)r_ * c_ * len_, 0u) {}
    uint32_t* at(uint32_t i, uint32_t j) { return a.data() + ((uint64_t)i * c + j) * len; }
    const uint32_t* at(uint32_t i, uint32_t j) const {
        return a.data() + ((uint64_t)i * c + j) * len;
    }
};

PolyMat identity(uint32_t n) {
    PolyMat I(n, n, 1);
    for (uint32_t i = 0; i < n; ++i) I.at(i, i)[0] = 1;
    return I;
}

// drops trailing coefficients that are zero in every entry (len stays >= 1)
void trim(PolyMat& M) {
    uint64_t top = 1;
    for (uint64_t e = 0; e < (uint64_t)M.r * M.c; ++e) {
        const uint32_t* x = M.a.data() + e * M.len;
        for (uint64_t k = M.len; k > top; --k)
            if (x[k - 1]) {
                top = k;
                break;
            }
    }
    if (top == M.len) return;
    PolyMat T(M.r, M.c, top);
    for (uint64_t e = 0; e < (uint64_t)M.r * M.c; ++e)
        std::memcpy(T.a.data() + e * top, M.a.data() + e * M.len, top * sizeof(uint32_t));
    M = std::move(T);
}

bool entry_nonzero(const uint32_t* x, uint64_t n) {
    for (uint64_t k = 0; k < n; ++k)
        if (x[k]) return true;
    return false;
}

// ---------------------------------------------------------------------------
// NTT over word-size primes q < 2^31 (Montgomery form, R = 2^32)
// ---------------------------------------------------------------------------
struct NttField {
    uint32_t q = 0, g = 0, lg_max = 0;
    uint32_t qneg_inv = 0;  // -q^-1 mod 2^32
    uint32_t r2 = 0;        // 2^64 mod q

    NttField(uint32_t q_, uint32_t g_, uint32_t lg_) : q(q_), g(g_), lg_max(lg_) {
        uint32_t inv = q;  // Newton iteration for q^-1 mod 2^32 (q odd)
        for (int i = 0; i < 5; ++i) inv *= 2u - q * inv;
        qneg_inv = 0u - inv;
        r2 = (uint32_t)(((unsigned __int128)1 << 64) % q);
    }
    // a * b * 2^-32 mod q for a, b < q
    uint32_t mul(uint32_t a, uint32_t b) const {
        const uint64_t t = (uint64_t)a * b;
        const uint32_t m = (uint32_t)t * qneg_inv;
        const uint64_t u = (t + (uint64_t)m * q) >> 32;
        return u >= q ? (uint32_t)(u - q) : (uint32_t)u;
    }
    uint32_t add(uint32_t a, uint32_t b) const {
        const uint32_t s = a + b;
        return s >= q ? s - q : s;
    }
    uint32_t sub(uint32_t a, uint32_t b) const { return a >= b ? a - b : a + q - b; }
    uint32_t to_mont(uint32_t a) const { return mul(a, r2); }
    uint32_t pow_plain(uint32_t a, uint64_t e) const {
        uint64_t r = 1, x = a % q;
        while (e) {
            if (e & 1) r = r * x % q;
            x = x * x % q;
            e >>= 1;
        }
        return (uint32_t)r;
    }
};

// Primes c * 2^k + 1 with a primitive root; the product of the first three
// exceeds 2^90, of all four 2^119.
const NttField kNttFields[4] = {
    NttField(2013265921u, 31, 27),  // 15 * 2^27 + 1
    NttField(1811939329u, 13, 26),  // 27 * 2^26 + 1
    NttField(469762049u, 3, 26),    //  7 * 2^26 + 1
    NttField(754974721u, 11, 24),   // 45 * 2^24 + 1
};

struct NttPlan {
    const NttField* f = nullptr;
    uint64_t n = 0;
    std::vector<uint32_t> fw, iv;  // level tables: [len + j] = w_(2 len)^(+-j), Montgomery
    uint32_t n_inv = 0;            // n^-1 mod q, plain: multiplying by it leaves Montgomery form

    NttPlan(const NttField& field, uint32_t lg) : f(&field), n(1ull << lg), fw(n), iv(n) {
        for (uint64_t len = 1; len < n; len <<= 1) {
            const uint32_t w = f->pow_plain(f->g, (f->q - 1) / (2 * len));
            const uint32_t wi = f->pow_plain(w, f->q - 2);
            const uint32_t wm = f->to_mont(w), wim = f->to_mont(wi);
            uint32_t x = f->to_mont(1), y = x;
            for (uint64_t j = 0; j < len; ++j) {
                fw[len + j] = x;
                iv[len + j] = y;
                x = f->mul(x, wm);
                y = f->mul(y, wim);
            }
        }
        n_inv = f->pow_plain((uint32_t)(n % f->q), f->q - 2);
    }
    // natural order in, bit-reversed out (Gentleman-Sande)
    void forward(uint32_t* x) const {
        for (uint64_t len = n >> 1; len >= 1; len >>= 1) {
            const uint32_t* w = fw.data() + len;
            for (uint64_t i = 0; i < n; i += 2 * len)
                for (uint64_t j = 0; j < len; ++j) {
                    const uint32_t u = x[i + j], v = x[i + j + len];
                    x[i + j] = f->add(u, v);
                    x[i + j + len] = f->mul(f->sub(u, v), w[j]);
                }
        }
    }
    // bit-reversed in, natural order out, times n (Cooley-Tukey)
    void inverse(uint32_t* x) const {
        for (uint64_t len = 1; len < n; len <<= 1) {
            const uint32_t* w = iv.data() + len;
            for (uint64_t i = 0; i < n; i += 2 * len)
                for (uint64_t j = 0; j < len; ++j) {
                    const uint32_t u = x[i + j], v = f->mul(x[i + j + len], w[j]);
                    x[i + j] = f->add(u, v);
                    x[i + j + len] = f->sub(u, v);
                }
        }
    }
};

constexpr uint64_t kSchoolbookLen = 32;  // shorter operand at or below this: schoolbook

// Coefficients [lo, hi) of A(x) B(x) mod p (A: r x s, B: s x c), as an
// r x c matrix of length hi - lo.  With the NTT, a cyclic length N >= hi with
// N >= la + lb - 1 - lo suffices: wrapped terms land below lo (middle product).
PolyMat mul_range(const PolyMat& A, const PolyMat& B, uint64_t lo, uint64_t hi, const ModP& m,
                  uint32_t threads) {
    require(A.c == B.r, "polynomial matrix product: inner dimensions differ");
    PolyMat C(A.r, B.c, hi - lo);
    const uint64_t la = std::min(A.len, hi), lb = std::min(B.len, hi);
    if (la == 0 || lb == 0 || lo >= la + lb - 1) return C;
    const uint64_t top = std::min(hi, la + lb - 1);  // outputs at or above are zero
    const uint32_t R = A.r, S = A.c, Cn = B.c;
    std::vector<char> nzA((uint64_t)R * S), nzB((uint64_t)S * Cn);
    for (uint32_t i = 0; i < R; ++i)
        for (uint32_t t = 0; t < S; ++t) nzA[(uint64_t)i * S + t] = entry_nonzero(A.at(i, t), la);
    for (uint32_t t = 0; t < S; ++t)
        for (uint32_t j = 0; j < Cn; ++j) nzB[(uint64_t)t * Cn + j] = entry_nonzero(B.at(t, j), lb);

    if (std::min(la, lb) <= kSchoolbookLen) {
        // an output coefficient sums at most S * min(la, lb) products; when
        // that many (p-1)^2 fit in 64 bits the products are added unreduced
        const bool raw = (unsigned __int128)(m.p - 1) * (m.p - 1) * S * std::min(la, lb) <
                         ((unsigned __int128)1 << 64);
        parallel_for((uint64_t)R * Cn, threads, [&](uint64_t e) {
            const uint32_t i = (uint32_t)(e / Cn), j = (uint32_t)(e % Cn);
            std::vector<uint64_t> acc(top - lo, 0);  // sums of residues < 2^31: no overflow
            for (uint32_t t = 0; t < S; ++t) {
                if (!nzA[(uint64_t)i * S + t] || !nzB[(uint64_t)t * Cn + j]) continue;
                const uint32_t* a = A.at(i, t);
                const uint32_t* b = B.at(t, j);
                for (uint64_t ka = 0; ka < la && ka < top; ++ka) {
                    if (!a[ka]) continue;
                    const uint64_t kb0 = lo > ka ? lo - ka : 0;
                    const uint64_t kb1 = std::min(lb, top - ka);
                    if (kb0 >= kb1) continue;
                    uint64_t* dst = acc.data() + (ka + kb0 - lo);
                    const uint32_t* src = b + kb0;
                    const uint64_t cnt = kb1 - kb0;
                    const uint64_t av = a[ka];
                    if (raw) {
                        for (uint64_t q = 0; q < cnt; ++q) dst[q] += av * src[q];
                    } else {
                        for (uint64_t q = 0; q < cnt; ++q) dst[q] += mulmod(a[ka], src[q], m);
                    }
                }
            }
            uint32_t* out = C.at(i, j);
            for (uint64_t k = lo; k < top; ++k) out[k - lo] = mod_u64(acc[k - lo], m);
        });
        return C;
    }

    // transform length and number of primes (the integer sums stay below Q)
    const uint64_t need = std::max(top, la + lb - 1 - lo);
    uint32_t lg = 0;
    while ((1ull << lg) < need) ++lg;
    const unsigned __int128 bound =
        (unsigned __int128)S * std::min(la, lb) * (uint64_t)(m.p - 1) * (uint64_t)(m.p - 1);
    uint32_t np = 0;
    unsigned __int128 Q = 1;
    while (np < 4 && Q <= bound) Q *= kNttFields[np++].q;
    require(Q > bound, "polynomial product exceeds the CRT range of four NTT primes");
    for (uint32_t q = 0; q < np; ++q)
        require(lg <= kNttFields[q].lg_max,
                "polynomial product longer than the NTT primes support (2^24 coefficients)");
    const uint64_t n = 1ull << lg, w = top - lo;

    // residues per prime, output-major [prime][entry][k - lo]
    std::vector<std::vector<uint32_t>> res(np, std::vector<uint32_t>((uint64_t)R * Cn * w, 0));
    for (uint32_t q = 0; q < np; ++q) {
        const NttField& f = kNttFields[q];
        const NttPlan plan(f, lg);
        std::vector<uint32_t> FA((uint64_t)R * S * n), FB((uint64_t)S * Cn * n);
        parallel_for((uint64_t)R * S, threads, [&](uint64_t e) {
            if (!nzA[e]) return;
            uint32_t* x = FA.data() + e * n;
            const uint32_t* src = A.a.data() + e * A.len;
            for (uint64_t k = 0; k < la; ++k) x[k] = f.to_mont(src[k] % f.q);
            std::fill(x + la, x + n, 0u);
            plan.forward(x);
        });
        parallel_for((uint64_t)S * Cn, threads, [&](uint64_t e) {
            if (!nzB[e]) return;
            uint32_t* x = FB.data() + e * n;
            const uint32_t* src = B.a.data() + e * B.len;
            for (uint64_t k = 0; k < lb; ++k) x[k] = f.to_mont(src[k] % f.q);
            std::fill(x + lb, x + n, 0u);
            plan.forward(x);
        });
        parallel_for((uint64_t)R * Cn, threads, [&](uint64_t e) {
            const uint32_t i = (uint32_t)(e / Cn), j = (uint32_t)(e % Cn);
            std::vector<uint32_t> acc(n, 0);
            bool any = false;
            for (uint32_t t = 0; t < S; ++t) {
                const uint64_t ea = (uint64_t)i * S + t, eb = (uint64_t)t * Cn + j;
                if (!nzA[ea] || !nzB[eb]) continue;
                any = true;
                const uint32_t* x = FA.data() + ea * n;
                const uint32_t* y = FB.data() + eb * n;
                for (uint64_t k = 0; k < n; ++k) acc[k] = f.add(acc[k], f.mul(x[k], y[k]));
            }
            if (!any) return;
            plan.inverse(acc.data());
            uint32_t* out = res[q].data() + e * w;
            for (uint64_t k = lo; k < top; ++k) out[k - lo] = f.mul(acc[k], plan.n_inv);
        });
    }

    // Garner: x = v0 + v1 q0 + v2 q0 q1 + ...  (0 <= x < Q), then x mod p
    uint32_t inv_tab[4][4] = {};
    uint32_t radix_p[4] = {};
    {
        uint64_t pr = 1 % m.p;
        for (uint32_t i = 0; i < np; ++i) {
            radix_p[i] = (uint32_t)pr;
            pr = mulmod((uint32_t)pr, kNttFields[i].q % m.p, m);
            for (uint32_t j = 0; j < i; ++j)
                inv_tab[i][j] = kNttFields[i].pow_plain(kNttFields[j].q % kNttFields[i].q,
                                                        kNttFields[i].q - 2);
        }
    }
    parallel_for((uint64_t)R * Cn, threads, [&](uint64_t e) {
        uint32_t* out = C.a.data() + e * C.len;
        for (uint64_t k = 0; k < w; ++k) {
            uint64_t v[4];
            uint64_t acc = 0;
            for (uint32_t i = 0; i < np; ++i) {
                const uint64_t qi = kNttFields[i].q;
                uint64_t t = res[i][e * w + k];
                for (uint32_t j = 0; j < i; ++j) t = (t + qi - v[j] % qi) % qi * inv_tab[i][j] % qi;
                v[i] = t;
                acc += mulmod((uint32_t)(t % m.p), radix_p[i], m);
            }
            out[k] = mod_u64(acc, m);
        }
    });
    return C;
}

// ---------------------------------------------------------------------------
// order bases
// ---------------------------------------------------------------------------

// dst -= f * src over [from, to)
void sub_scaled(uint32_t* dst, const uint32_t* src, uint64_t from, uint64_t to, uint32_t f,
                const ModP& m) {
    const uint32_t nf = f ? m.p - f : 0;
    if (m.p <= 65536u) {
        // nf * src < 2^32: one 32-bit Barrett step with mu = floor(2^32 / p)
        // leaves t - qt p in [0, 3p); two conditional subtractions finish it
        // (vectorisable, no 128-bit products)
        const uint32_t p = m.p;
        const uint64_t mu = ((uint64_t)1 << 32) / p;
        for (uint64_t q = from; q < to; ++q) {
            const uint32_t t = nf * src[q];
            uint32_t r = t - (uint32_t)(((uint64_t)t * mu) >> 32) * p;
            r = r >= p ? r - p : r;
            r = r >= p ? r - p : r;
            const uint32_t x = dst[q] + r;  // < 2p <= 2^17
            dst[q] = x >= p ? x - p : x;
        }
        return;
    }
    for (uint64_t q = from; q < to; ++q) dst[q] = mod_u64((uint64_t)dst[q] + (uint64_t)nf * src[q], m);
}

// d elimination steps on the residual G (m x n, coefficients [0, d)): the
// reference's `basis` (sigma_basis.rs:233-309) applied d times, each to the
// next coefficient of the updated residual.  Returns the m x m basis; delta
// is updated (rows of the result follow the final sorted order).
PolyMat mbasis(const PolyMat& G, uint64_t d, std::vector<uint64_t>& delta, const ModP& mp) {
    const uint32_t m = G.r, n = G.c;
    std::vector<std::vector<uint32_t>> R(m, std::vector<uint32_t>(d * n, 0));  // [k][col]
    for (uint32_t i = 0; i < m; ++i)
        for (uint32_t j = 0; j < n; ++j) {
            const uint32_t* g = G.at(i, j);
            for (uint64_t k = 0; k < d && k < G.len; ++k) R[i][k * n + j] = g[k];
        }
    std::vector<std::vector<uint32_t>> M(m);  // [degree][col]
    std::vector<uint64_t> mlen(m, 1);
    for (uint32_t i = 0; i < m; ++i) {
        M[i].assign(m, 0);
        M[i][i] = 1;
    }
    std::vector<uint32_t> perm(m), pos(n);
    std::vector<char> xrow(m);
    struct Pivot {
        uint32_t row, col, inv;
    };
    std::vector<Pivot> piv;
    for (uint64_t k = 0; k < d; ++k) {
        // stable sort by delta (Rust's sort_by is stable)
        std::iota(perm.begin(), perm.end(), 0u);
        std::stable_sort(perm.begin(), perm.end(),
                         [&](uint32_t x, uint32_t y) { return delta[x] < delta[y]; });
        {
            std::vector<uint64_t> nd(m), nl(m);
            std::vector<std::vector<uint32_t>> nR(m), nM(m);
            for (uint32_t q = 0; q < m; ++q) {
                nd[q] = delta[perm[q]];
                nl[q] = mlen[perm[q]];
                nR[q].swap(R[perm[q]]);
                nM[q].swap(M[perm[q]]);
            }
            delta.swap(nd);
            mlen.swap(nl);
            R.swap(nR);
            M.swap(nM);
        }
        // LSP elimination of coefficient k (modular_linalg.rs:71-125): rows in
        // order; each row is reduced by the earlier pivot rows; its pivot is
        // the first nonzero at column positions >= rank of the running column
        // permutation, whose entries at `rank` and the pivot are then swapped
        std::iota(pos.begin(), pos.end(), 0u);
        uint32_t rank = 0;
        piv.clear();
        std::fill(xrow.begin(), xrow.end(), 0);
        for (uint32_t i = 0; i < m; ++i) {
            uint32_t* Ri = R[i].data();
            for (const Pivot& pv : piv) {
                const uint32_t v = Ri[k * n + pv.col];
                if (!v) continue;
                const uint32_t f = mulmod(v, pv.inv, mp);
                sub_scaled(Ri, R[pv.row].data(), k * n, d * n, f, mp);
                if (mlen[pv.row] > mlen[i]) {
                    M[i].resize(mlen[pv.row] * m, 0);
                    mlen[i] = mlen[pv.row];
                }
                sub_scaled(M[i].data(), M[pv.row].data(), 0, mlen[pv.row] * m, f, mp);
            }
            for (uint32_t jj = rank; jj < n; ++jj) {
                const uint32_t c = pos[jj];
                if (!Ri[k * n + c]) continue;
                std::swap(pos[rank], pos[jj]);
                ++rank;
                piv.push_back({i, c, inv_mod(Ri[k * n + c], mp)});
                xrow[i] = 1;
                break;
            }
        }
        // pivot rows times x (delta + 1); the others keep their degree
        for (uint32_t i = 0; i < m; ++i) {
            if (!xrow[i]) continue;
            std::vector<uint32_t>& Ri = R[i];
            std::memmove(Ri.data() + (k + 1) * n, Ri.data() + k * n, (d - k - 1) * n * sizeof(uint32_t));
            std::fill(Ri.begin() + k * n, Ri.begin() + (k + 1) * n, 0u);
            M[i].insert(M[i].begin(), m, 0u);
            mlen[i] += 1;
            delta[i] += 1;
        }
    }
    uint64_t len = 1;
    for (uint32_t i = 0; i < m; ++i) len = std::max(len, mlen[i]);
    PolyMat out(m, m, len);
    for (uint32_t i = 0; i < m; ++i)
        for (uint64_t deg = 0; deg < mlen[i]; ++deg)
            for (uint32_t j = 0; j < m; ++j) out.at(i, j)[deg] = M[i][deg * m + j];
    trim(out);
    return out;
}

constexpr uint64_t kLeafOrder = 16;  // orders at or below: iterative steps

// pm_basis (sigma_basis.rs:37-73): M' for the first d/2 coefficients, the
// residual x^-(d/2) M' G mod x^(d - d/2), M'' for it, M = M'' M'.
PolyMat pm_basis(const PolyMat& G, uint64_t d, std::vector<uint64_t>& delta, const ModP& mp,
                 uint32_t threads) {
    if (d == 0) return identity(G.r);
    if (d <= kLeafOrder) return mbasis(G, d, delta, mp);
    const uint64_t d1 = d / 2;
    PolyMat M1 = pm_basis(G, d1, delta, mp, threads);
    PolyMat G2 = mul_range(M1, G, d1, d, mp, threads);
    PolyMat M2 = pm_basis(G2, d - d1, delta, mp, threads);
    PolyMat M = mul_range(M2, M1, 0, M2.len + M1.len - 1, mp, threads);
    trim(M);
    return M;
}

}  // namespace

// RankComputation (third_party/bcw/src/rank_pipeline.rs:105-287)
struct pgw_rank {
    uint32_t b = 0;
    uint64_t n_rows = 0, n_cols = 0;
    uint32_t threshold = 0, threads = 1;
    uint64_t max_nlen = 0;
    ModP m{};
    std::vector<uint64_t> delta;
    PolyMat basis;
    uint64_t processed = 0, spread = 0, rank = 0;

    int status() const {
        if (spread >= threshold) return PGW_RANK_SUCCESS;
        if (rank > std::min(n_rows, n_cols)) return PGW_RANK_EXCEEDS_DIMENSION;
        if (processed >= max_nlen) return PGW_RANK_SEQUENCE_LIMIT;
        return PGW_RANK_RUNNING;
    }
    void fill(pgw_rank_info* info) const {
        info->rank = rank;
        info->delta_spread = spread;
        info->processed = processed;
        info->max_terms = max_nlen;
        info->status = status();
    }
};

extern "C" int pgw_rank_create(uint32_t prime, uint32_t block, uint64_t n_rows, uint64_t n_cols,
                               uint32_t threshold, uint32_t max_nlen_buffer, uint32_t threads,
                               pgw_rank** out) {
    return guarded([&] {
        require(out != nullptr, "null output");
        *out = nullptr;
        require(prime >= 2 && prime < (1u << 31), "prime must satisfy 2 <= p < 2^31");
        for (uint32_t q = 2; (uint64_t)q * q <= prime; ++q)
            require(prime % q != 0, "modulus is not prime");
        require(block >= 1 && block <= 64, "block width must be in 1..64");
        require(n_rows > 0 && n_cols > 0, "matrix dimensions must be nonzero");
        auto r = std::make_unique<pgw_rank>();
        r->b = block;
        r->n_rows = n_rows;
        r->n_cols = n_cols;
        r->threshold = threshold;
        if (threads == 0) threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        r->threads = threads;
        // max_nlen (rank_pipeline.rs:312-314)
        r->max_nlen = 2 * std::min(n_rows, n_cols) / block + max_nlen_buffer;
        r->m = make_modp(prime);
        r->delta.assign(2 * block, 0);
        for (uint32_t i = block; i < 2 * block; ++i) r->delta[i] = 1;
        r->basis = identity(2 * block);
        *out = r.release();
    });
}

extern "C" int pgw_rank_destroy(pgw_rank* r) {
    delete r;
    return PGW_OK;
}

extern "C" int pgw_rank_step(pgw_rank* r, const uint32_t* h_seq, uint64_t len, pgw_rank_info* info) {
    return guarded([&] {
        require(r && info, "null rank state or output");
        require(len == 0 || h_seq, "null sequence");
        const uint32_t b = r->b;
        const uint64_t ut = (uint64_t)b * (b + 1) / 2;
        const uint64_t usable = len ? len - 1 : 0;  // the initial V^T V term is dropped
        if (usable <= r->processed)
            throw Error(PGW_ERR_NEED_MORE_SEQUENCE,
                        "need more sequence: " + std::to_string(usable) + " usable terms, " +
                            std::to_string(r->processed) + " already processed");
        uint32_t lo, hi;
        min_max_u32(h_seq, len * ut, lo, hi);
        require(hi < r->m.p, "sequence entries must be reduced modulo p");
        (void)lo;
        // G = [S(x); I_b] (process_input_sequence, sigma_basis.rs:75-110)
        PolyMat G(2 * b, b, usable);
        for (uint64_t t = 0; t < usable; ++t) {
            const uint32_t* s = h_seq + (t + 1) * ut;
            uint64_t e = 0;
            for (uint32_t i = 0; i < b; ++i)
                for (uint32_t j = i; j < b; ++j, ++e) {
                    G.at(i, j)[t] = s[e];
                    G.at(j, i)[t] = s[e];
                }
        }
        for (uint32_t i = 0; i < b; ++i) G.at(b + i, i)[0] = 1;
        // residual of the new terms under the basis so far, then its order basis
        const uint64_t done = r->processed;
        PolyMat shifted = mul_range(r->basis, G, done, usable, r->m, r->threads);
        std::vector<uint64_t> delta = r->delta;
        PolyMat update = pm_basis(shifted, usable - done, delta, r->m, r->threads);
        PolyMat basis = mul_range(update, r->basis, 0, update.len + r->basis.len - 1, r->m, r->threads);
        trim(basis);
        // analyze_delta (sigma_basis.rs:312-324)
        std::vector<uint64_t> sorted = delta;
        std::sort(sorted.begin(), sorted.end());
        r->spread = sorted[b] - sorted[b - 1];
        r->rank = std::accumulate(sorted.begin(), sorted.begin() + b, (uint64_t)0);
        r->delta.swap(delta);
        r->basis = std::move(basis);
        r->processed = usable;
        r->fill(info);
    });
}

extern "C" int pgw_rank_get_info(const pgw_rank* r, pgw_rank_info* info) {
    return guarded([&] {
        require(r && info, "null rank state or output");
        r->fill(info);
    });
}

extern "C" int pgw_rank_delta(const pgw_rank* r, uint64_t* h_delta) {
    return guarded([&] {
        require(r && h_delta, "null rank state or output");
        std::copy(r->delta.begin(), r->delta.end(), h_delta);
    });
}

// order_basis_to_generator_matrixlist + save_generator_list
// (sigma_basis.rs:112-168): line k = rev_(i,j)[k] for i, j < b row-major with
// rev_(i,j)[k] = basis[i][j][delta_i - k]; lines after the last nonzero dropped
extern "C" int pgw_rank_generator(const pgw_rank* r, uint32_t* h_out, uint64_t capacity,
                                  uint64_t* terms) {
    return guarded([&] {
        require(r != nullptr && terms != nullptr, "null rank state or output");
        const uint32_t b = r->b;
        uint64_t maxdel = 0;
        for (uint32_t i = 0; i < b; ++i) maxdel = std::max(maxdel, r->delta[i]);
        uint64_t maxdeg = 0;
        for (uint32_t i = 0; i < b; ++i)
            for (uint32_t j = 0; j < b; ++j) {
                const uint32_t* x = r->basis.at(i, j);
                for (uint64_t k = 1; k <= r->delta[i]; ++k) {
                    const uint64_t src = r->delta[i] - k;
                    if (src < r->basis.len && x[src] && k > maxdeg) maxdeg = k;
                }
            }
        *terms = maxdeg + 1;
        if (!h_out) return;
        require(capacity >= (maxdeg + 1) * b * b,
                "generator buffer too small: " + std::to_string((maxdeg + 1) * b * b) +
                    " values needed");
        for (uint64_t k = 0; k <= maxdeg; ++k)
            for (uint32_t i = 0; i < b; ++i)
                for (uint32_t j = 0; j < b; ++j) {
                    const uint64_t del = r->delta[i];
                    uint32_t v = 0;
                    if (k <= del && del - k < r->basis.len) v = r->basis.at(i, j)[del - k];
                    h_out[(k * b + i) * b + j] = v;
                }
    });
}

"""Pins of the CPU oracle (oracle/pgw_oracle.c) against facts it does not compute itself.

Each test checks the oracle with something independent of its code: published
test vectors, the golden fixtures of the reference's own tests, brute-force
dense assembly from the mathematical definition, or invariants.
"""
import itertools
import json
from math import comb

import numpy as np
import pytest

import oracle_ctypes as oc

GOLDEN = oc.ROOT / "tests" / "golden"


# ---------------------------------------------------------------- subsets ----

@pytest.mark.parametrize("n,k", [(5, 0), (5, 2), (7, 3), (9, 4), (12, 5), (10, 10)])
def test_unrank_matches_itertools_lex_order(n, k):
    L = oc.lib()
    assert L.orc_binom(n, k) == comb(n, k)
    out = np.zeros(max(k, 1), np.uint32)
    for rank, subset in enumerate(itertools.combinations(range(n), k)):
        assert L.orc_unrank_subset(rank, n, k, out) == 0
        assert tuple(out[:k]) == subset
    assert L.orc_unrank_subset(comb(n, k), n, k, out) == -1


# ------------------------------------------------------- RNG and hashes ----

def test_splitmix64_published_vector():
    # Vigna's splitmix64 from state 0
    st = oc.ctypes.c_uint64(0)
    got = [oc.lib().orc_splitmix64_next(oc.ctypes.byref(st)) for _ in range(3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_fnv1a64_published_vectors():
    L = oc.lib()
    assert L.orc_fnv1a64_string(b"") == 0xCBF29CE484222325
    assert L.orc_fnv1a64_string(b"a") == 0xAF63DC4C8601EC8C
    assert L.orc_fnv1a64_string(b"foobar") == 0x85944171F73967E8


def test_mt19937_64_standard_value():
    # [rand.predef]: the 10000th output of a default-constructed mt19937_64
    out = np.zeros(10000, np.uint64)
    oc.lib().orc_mt19937_64_outputs(5489, 10000, out)
    assert int(out[-1]) == 9981545732273789042


def test_uniform_int_matches_libstdcxx_golden():
    g = json.loads((GOLDEN / "libstdcxx_uniform.json").read_text())
    assert g["mt19937_64_default_10000th"] == 9981545732273789042
    for case in g["cases"]:
        out = np.zeros(len(case["values"]), np.uint64)
        oc.lib().orc_uniform_u64(case["seed"], case["lo"], case["hi"], out.size, out)
        assert out.tolist() == case["values"], case["seed"]


def test_reference_random_state_draw_order():
    # Dr (nonzero), Dc (nonzero), V from one stream: the concatenation equals
    # the individual uniform draws of the same stream
    p, rows, cols, b = 65521, 7, 5, 3
    seed = oc.lib().orc_seeded_sector_value(1, 2)
    dr, dc, v = oc.reference_random_state(1, 2, p, rows, cols, b)
    g = json.loads((GOLDEN / "libstdcxx_uniform.json").read_text())  # presence only
    assert g
    nz = np.zeros(rows + cols, np.uint64)
    oc.lib().orc_uniform_u64(seed, 1, p - 1, rows + cols, nz)
    assert dr.tolist() == nz[:rows].tolist()
    assert dc.tolist() == nz[rows:].tolist()
    assert (v < p).all() and (dr > 0).all() and (dc > 0).all()


# ------------------------------------------------------------- row mixing --

def _dense_rowmix(pairs, rows, rounds, p):
    P = np.eye(rows, dtype=object)
    pc = rows // 2
    for r in range(rounds):
        Pr = np.eye(rows, dtype=object)
        for f, s, a, b in pairs.reshape(rounds, pc, 4)[r].tolist():
            Pr[f, f], Pr[f, s], Pr[s, f], Pr[s, s] = 1, a, b, (1 + a * b)
        P = Pr.dot(P)
    return P


def test_rowmix_plan_is_two_rounds_of_disjoint_pairs():
    p, rows = 65521, 101
    realized, pairs, _ = oc.rowmix_plan(10001, "eliminated", 3, rows, p, 2)
    pr = pairs.reshape(2, rows // 2, 4)
    for r in range(2):
        touched = pr[r, :, :2].reshape(-1)
        assert len(set(touched.tolist())) == touched.size  # disjoint
        assert (pr[r, :, 2:] >= 1).all() and (pr[r, :, 2:] < p).all()


# ---------------------------------------------------- synthetic operator --

def _dense_synthetic(n, k, M, Mp, T, E, p):
    """Dense D (R x N) and E (N x R) straight from the exterior-algebra definition."""
    src = list(itertools.combinations(range(n), k))
    tgt = list(itertools.combinations(range(n), k - 1))
    tidx = {s: i for i, s in enumerate(tgt)}
    T = np.asarray(T, dtype=object).reshape(n, Mp, M)
    E = np.asarray(E, dtype=object).reshape(n, Mp, M)
    D = np.zeros((len(tgt) * Mp, len(src) * M), dtype=object)
    Em = np.zeros((len(src) * M, len(tgt) * Mp), dtype=object)
    for c, s in enumerate(src):
        for pos, i in enumerate(s):
            row = tidx[s[:pos] + s[pos + 1:]]
            sign = -1 if pos % 2 else 1
            for beta in range(Mp):
                for a in range(M):
                    D[row * Mp + beta, c * M + a] = sign * T[i, beta, a]
                    Em[c * M + a, row * Mp + beta] = sign * E[i, beta, a]
    return D % p, Em % p


@pytest.mark.parametrize("n,k,M,Mp,p,b,rounds", [(5, 2, 3, 4, 101, 2, 2), (6, 3, 2, 3, 65521, 3, 1),
                                                  (4, 1, 3, 2, 2147483647, 2, 2), (5, 5, 2, 3, 97, 1, 0)])
def test_synthetic_operator_matches_dense_assembly(n, k, M, Mp, p, b, rounds):
    rng = np.random.default_rng(n * 100 + k)
    T = oc.synthetic_table(7, p, n * Mp * M)
    E = oc.synthetic_table(8, p, n * Mp * M)
    N = comb(n, k) * M
    dr, dc, _ = oc.reference_random_state(1, 0, p, N, N, b)
    realized, pairs, _ = oc.rowmix_plan(10001, "synthetic", 0, N, p, rounds)
    op = oc.synthetic_op(n, k, M, Mp, p, T, E, dr, dc, rounds, pairs if rounds else None)
    assert (op.rows, op.cols) == (N, N)
    D, Em = _dense_synthetic(n, k, M, Mp, T, E, p)
    A = Em.dot(D) % p
    X = rng.integers(0, p, size=(N, b)).astype(object)
    Dc = np.diag(dc.astype(object))
    Dr = np.diag(dr.astype(object))
    fwd = A.dot(Dc).dot(X) % p
    got = op.apply_forward(np.asarray(X, dtype=np.uint64).astype(np.uint32), b).reshape(N, b)
    assert (got.astype(object) == fwd).all()
    P = _dense_rowmix(pairs, N, rounds, p) if rounds else np.eye(N, dtype=object)
    S = Dc.dot(A.T).dot(P.T).dot(Dr).dot(P).dot(A).dot(Dc) % p
    want = S.dot(X) % p
    got = op.apply_square(np.asarray(X, dtype=np.uint64).astype(np.uint32), b).reshape(N, b)
    assert (got.astype(object) == want).all()


def test_sequence_closed_form_against_dense_powers():
    n, k, M, Mp, p, b = 5, 2, 2, 3, 10007, 2
    T = oc.synthetic_table(3, p, n * Mp * M)
    E = oc.synthetic_table(4, p, n * Mp * M)
    N = comb(n, k) * M
    dr, dc, v = oc.reference_random_state(5, 0, p, N, N, b)
    realized, pairs, _ = oc.rowmix_plan(10001, "synthetic", 0, N, p, 2)
    op = oc.synthetic_op(n, k, M, Mp, p, T, E, dr, dc, 2, pairs)
    D, Em = _dense_synthetic(n, k, M, Mp, T, E, p)
    A = Em.dot(D) % p
    P = _dense_rowmix(pairs, N, 2, p)
    Dc, Dr = np.diag(dc.astype(object)), np.diag(dr.astype(object))
    S = Dc.dot(A.T).dot(P.T).dot(Dr).dot(P).dot(A).dot(Dc) % p
    seq, curv = op.sequence(v, b, 11)  # L = 13
    assert seq.shape == (13, 3)
    V = v.reshape(N, b).astype(object)
    W = V.copy()
    powers = [V]
    for t in range(13):
        G = V.T.dot(W) % p
        assert seq[t].tolist() == [G[0, 0], G[0, 1], G[1, 1]], t
        W = S.dot(W) % p
        powers.append(W)
    # two applies per round of four terms: the saved block is S^(2*rounds) V
    assert (curv.reshape(N, b).astype(object) == powers[6]).all()


def test_gram_upper_packing_order():
    p, b, rows = 65521, 4, 37
    rng = np.random.default_rng(1)
    x = rng.integers(0, p, (rows, b)).astype(np.uint32)
    y = rng.integers(0, p, (rows, b)).astype(np.uint32)
    got = oc.gram_upper(x, y, rows, b, p)
    G = x.astype(object).T.dot(y.astype(object)) % p
    want = [G[i, j] for i in range(b) for j in range(i, b)]
    assert got.tolist() == want


# ------------------------------------------------ reference cyclic operator --

def _dense_eliminated(inst):
    """F = [-D2 R Dm | D2 B] over U = V minus w, global lexicographic order
    (x then z), from the differential's definition."""
    g, p = inst["genus"], inst["modulus"]
    n, m = g - 3, g // 2
    e = inst["third_section_elimination"]
    w = e["w_V_index"]
    mu = np.asarray(inst["artinian"]["mu"]["values"], dtype=np.int64).reshape(n, n, g)
    R = np.asarray(e["right_inverse_A0_by_A1"], dtype=np.int64).reshape(g, n)
    Bk = np.asarray(e["kernel_basis_A0_by_3"], dtype=np.int64).reshape(g, -1)
    U = [i for i in range(n) if i != w]
    sub = {d: list(itertools.combinations(U, d)) for d in (m, m - 1, m - 2)}

    def differential(src, tgt):
        ti = {s: i for i, s in enumerate(tgt)}
        Dmat = np.zeros((len(tgt) * n, len(src) * g), dtype=np.int64)
        for c, s in enumerate(src):
            for pos, i in enumerate(s):
                r = ti[s[:pos] + s[pos + 1:]]
                Dmat[r * n:(r + 1) * n, c * g:(c + 1) * g] += (-1) ** pos * mu[i]
        return Dmat % p

    Dm = differential(sub[m], sub[m - 1])
    D2 = differential(sub[m - 1], sub[m - 2])
    I = np.eye(len(sub[m - 1]), dtype=np.int64)
    Rb, Bb = np.kron(I, R), np.kron(I, Bk)
    x = (-(D2.dot(Rb) % p).dot(Dm)) % p
    z = D2.dot(Bb) % p
    return np.hstack([x, z]) % p


def _sector_dense(inst, sector):
    op = oc.cyclic_op(inst, sector)
    op.set_preconditioners(np.ones(op.rows, np.uint32), np.ones(op.cols, np.uint32))
    eye = np.eye(op.cols, dtype=np.uint32)
    return op, op.apply_forward(eye.reshape(-1), op.cols).reshape(op.rows, op.cols)


def test_g12_certificate_and_profiles(g12, g12_reference):
    p = g12["modulus"]
    ref = g12_reference
    total_nnz = 0
    for sector in range(g12["cyclic"]["order"]):
        op, F = _sector_dense(g12, sector)
        prof = ref["sector_profiles"][sector]
        assert (op.rows, op.cols) == (prof["rows"], prof["columns"])
        assert op.rows == g12["third_section_elimination"]["sector_rows"][sector]
        assert op.cols == g12["third_section_elimination"]["sector_columns"][sector]
        assert oc.rank_mod_p(F, p) == prof["rank"]
        total_nnz += int(np.count_nonzero(F))
        if sector == 0:
            K0 = np.asarray(ref["certificate"]["K0"], dtype=np.int64)
            assert not ((F.astype(np.int64) @ K0) % p).any()
            assert oc.rank_mod_p(K0.astype(np.uint32), p) == 2
    assert [ref["F_shape"][0], ref["F_shape"][1]] == [630, 504]
    assert total_nnz == ref["F_nnz_by_order"][0]


def test_g12_sectors_match_independent_dense_assembly(g12, g12_reference):
    F = _dense_eliminated(g12)
    assert F.shape == tuple(g12_reference["F_shape"])
    for sector in range(g12["cyclic"]["order"]):
        rows = g12_reference["certificate"]["target_sector_indices"][str(sector)]
        cols = g12_reference["certificate"]["source_sector_indices"][str(sector)]
        _, Fs = _sector_dense(g12, sector)
        assert (Fs.astype(np.int64) == F[np.ix_(rows, cols)]).all(), sector


def test_cyclic_transpose_identity_and_square(g12):
    p = g12["modulus"]
    op = oc.cyclic_op(g12, 3)
    rows, cols, b = op.rows, op.cols, 3
    dr, dc, v = oc.reference_random_state(1, 3, p, rows, cols, b)
    realized, pairs, _ = oc.rowmix_plan(10001, "eliminated", 3, rows, p, 2)
    op.set_preconditioners(dr, dc, 2, pairs)
    rng = np.random.default_rng(3)
    x = rng.integers(0, p, (cols, b)).astype(np.uint32)
    y = rng.integers(0, p, (rows, b)).astype(np.uint32)
    fx = op.apply_forward(x, b).astype(np.int64).reshape(rows, b)  # F Dc x
    fty = op.apply_backward(y, b).astype(np.int64).reshape(cols, b)  # Dc F^T y
    assert ((fx * y.astype(np.int64)).sum() - (x.astype(np.int64) * fty).sum()) % p == 0
    # S symmetric: <Sx, y'> = <x, Sy'>
    y2 = rng.integers(0, p, (cols, b)).astype(np.uint32)
    sx = op.apply_square(x, b).astype(np.int64).reshape(cols, b)
    sy = op.apply_square(y2, b).astype(np.int64).reshape(cols, b)
    assert ((sx * y2.astype(np.int64)).sum() - (x.astype(np.int64) * sy).sum()) % p == 0


def test_kernel_vectors_zero_certificate_on_exact_kernel(g12, g12_reference):
    # with one generator matrix G_0 = I, D = V; choosing V = Dc^-1 K0 gives F Dc D = F K0 = 0
    p = g12["modulus"]
    op = oc.cyclic_op(g12, 0)
    b = 2
    dc = np.arange(1, op.cols + 1, dtype=np.uint32) % (p - 1) + 1
    op.set_preconditioners(np.ones(op.rows, np.uint32), dc)
    K0 = np.asarray(g12_reference["certificate"]["K0"], dtype=np.int64)
    inv = np.array([pow(int(c), p - 2, p) for c in dc], dtype=np.int64)
    V = (K0 * inv[:, None]) % p
    gens = np.eye(b, dtype=np.uint32).reshape(-1)
    d, prod = op.kernel_vectors(V.astype(np.uint32), b, gens)
    assert (d.reshape(-1, b).astype(np.int64) == V).all()
    assert not prod.any()


# ------------------------------------------------------------------- WDM --

def test_wdm_writer_reproduces_bcw_fixture(tmp_path):
    # third_party/bcw/tests/complete_file.rs: A=diag(1,2), V=(1,1), p=29
    terms = 50
    seq, power = [], 1
    for _ in range(terms):
        seq.append((1 + power) % 29)
        power = 2 * power % 29
    expected = "2 2 29 50 1\n1 1\n1 1\n1 1\n1 1\n" + " ".join(map(str, seq)) + "\n"
    path = tmp_path / "case.wdm"
    one = np.ones(2, np.uint32)
    rc = oc.lib().orc_write_wdm(str(path).encode(), 2, 2, 29, one, one, one, one, 1,
                                np.asarray(seq, np.uint32), terms)
    assert rc == 0
    assert path.read_text() == expected

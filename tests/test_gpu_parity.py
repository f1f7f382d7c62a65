"""GPU parity: the HIP library against the CPU oracle, through the C ABI.

Bit-exact comparisons (integer arithmetic mod p, no tolerance) on seeded
inputs at sizes the oracle finishes in seconds, and size-independent
properties (two independent kernel paths agree; S is symmetric; linearity) at
the full single-sector sizes.
"""
import sys
from math import comb

import numpy as np
import pytest

import oracle_ctypes as oc
import prym_green_ancillary_amd as pg

pytestmark = pytest.mark.gpu

P31 = 2147483647
MFMA_CONFIGS = 10  # tile configurations of the MFMA sweep (mfma_config_count in the library)


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


@pytest.fixture(scope="module")
def ctx(torch_cuda):
    c = pg.Context(0)
    yield c
    c.close()


def dev(torch, a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.uint32))
    return torch.from_numpy(a.view(np.int32).copy()).cuda()


def host(t):
    return t.cpu().numpy().view(np.uint32).copy()


def stream(torch):
    return torch.cuda.current_stream().cuda_stream


def synthetic_setup(n, k, M, Mp, p, b, tseed=0, seed=1, rounds=2):
    spec = pg.SyntheticSpec(n, k, M, Mp, p, tseed, tseed + 1)
    T, E = spec.tables()
    N = comb(n, k) * M
    dr, dc, v = pg.reference_random_state(seed, 0, p, N, N, b)
    realized = pg.rowmix_realized_seed(10001, "synthetic", 0, N, p)
    pairs = pg.rowmix_plan(realized, N, rounds, p)
    return spec, T, E, N, dr, dc, v, pairs


def gpu_apply(torch, op, x):
    d_in = dev(torch, x)
    d_out = torch.empty_like(d_in)
    op.apply_square(d_in.data_ptr(), d_out.data_ptr(), stream(torch))
    torch.cuda.synchronize()
    return host(d_out).reshape(-1)


# ------------------------------------------------------------ cyclic (g12) --

def set_sell_form(monkeypatch, form):
    # sparse-map entry forms: "packed", one u32 (column | coefficient <<
    # column bits), which g12's sizes allow; "pairs", (column, coefficient);
    # "slice", per slice of 64 rows packed offsets from the slice's smallest
    # column (the form of g28 and larger); "slice4", the same with 4-bit
    # offsets, so that g12 has slices of both kinds
    if form == "pairs":
        monkeypatch.setenv("PGW_SELL_PACK", "0")
    elif form in ("slice", "slice4"):
        monkeypatch.setenv("PGW_SELL_PACK", "slice")
        if form == "slice4":
            monkeypatch.setenv("PGW_SELL_SLICE_BITS", "4")


@pytest.mark.parametrize("b,form,mix_order", [(4, "packed", True), (3, "packed", True), (4, "pairs", True),
                                              (4, "packed", False), (5, "pairs", False), (4, "slice", True),
                                              (4, "slice4", True), (3, "slice4", False)])
def test_g12_apply_square_all_sectors(torch_cuda, ctx, g12, b, form, mix_order, monkeypatch):
    # mix_order: t and M3 = P^T Dr P in the row-mixing order (the default);
    # PGW_CYCLIC_ORDER=0 keeps the natural row order
    set_sell_form(monkeypatch, form)
    if not mix_order:
        monkeypatch.setenv("PGW_CYCLIC_ORDER", "0")
    p = g12["modulus"]
    for sector in range(g12["cyclic"]["order"]):
        _log(f"g12 sector {sector} b={b}")
        o = oc.cyclic_op(g12, sector)
        dr, dc, v = pg.reference_random_state(1, sector, p, o.rows, o.cols, b)
        realized = pg.rowmix_realized_seed(10001, "eliminated", sector, o.rows, p)
        pairs = pg.rowmix_plan(realized, o.rows, 2, p)
        o.set_preconditioners(dr, dc, 2, pairs.reshape(-1))
        op = pg.Operator.cyclic(ctx, g12, sector, b, dr, dc, 2, pairs)
        assert (op.rows, op.cols) == (o.rows, o.cols)
        got = gpu_apply(torch_cuda, op, v)
        want = o.apply_square(v, b)
        assert (got == want).all(), sector


def test_g12_certificate_on_gpu(torch_cuda, ctx, g12, g12_reference):
    p = g12["modulus"]
    o = oc.cyclic_op(g12, 0)
    ones_r, ones_c = np.ones(o.rows, np.uint32), np.ones(o.cols, np.uint32)
    op = pg.Operator.cyclic(ctx, g12, 0, 2, ones_r, ones_c, 0, None)
    K0 = np.asarray(g12_reference["certificate"]["K0"], dtype=np.uint32)
    d_in = dev(torch_cuda, K0)
    d_out = torch_cuda.zeros(op.rows * 2, dtype=torch_cuda.int32, device="cuda")
    op.apply_forward(d_in.data_ptr(), d_out.data_ptr(), stream(torch_cuda))
    torch_cuda.cuda.synchronize()
    assert not host(d_out).any()
    # a non-kernel vector is not annihilated
    d_in = dev(torch_cuda, np.ones((op.cols, 2), np.uint32))
    op.apply_forward(d_in.data_ptr(), d_out.data_ptr(), stream(torch_cuda))
    torch_cuda.cuda.synchronize()
    assert host(d_out).any()


@pytest.mark.parametrize("mix_order,fused,form", [(True, True, "packed"), (False, True, "packed"),
                                                 (True, False, "packed"), (True, True, "slice4"),
                                                 (True, True, "pairs")])
def test_g12_sequence_and_kernel_vectors(torch_cuda, ctx, g12, mix_order, fused, form, monkeypatch):
    # fused: the round's block projections summed inside the last sparse map
    # (b = 4, p < 2^16), else by the projection kernels
    set_sell_form(monkeypatch, form)
    if not mix_order:
        monkeypatch.setenv("PGW_CYCLIC_ORDER", "0")
    if not fused:
        monkeypatch.setenv("PGW_CYCLIC_FUSED_GRAM", "0")
    p, sector, b = g12["modulus"], 1, 4
    o = oc.cyclic_op(g12, sector)
    dr, dc, v = pg.reference_random_state(1, sector, p, o.rows, o.cols, b)
    realized = pg.rowmix_realized_seed(10001, "eliminated", sector, o.rows, p)
    pairs = pg.rowmix_plan(realized, o.rows, 2, p)
    o.set_preconditioners(dr, dc, 2, pairs.reshape(-1))
    op = pg.Operator.cyclic(ctx, g12, sector, b, dr, dc, 2, pairs)
    max_len = 2 * o.cols // b + 50  # the reference's default length
    seq, curv = op.sequence(v, max_len)
    oseq, ocurv = o.sequence(v, b, max_len)
    assert (seq == oseq).all() and (curv == ocurv).all()
    rng = np.random.default_rng(5)
    gens = rng.integers(0, p, size=(9, b * b)).astype(np.uint32)
    d, prod = op.kernel_vectors(v, gens)
    od, oprod = o.kernel_vectors(v, b, gens)
    assert (d == od).all() and (prod == oprod).all()


# ----------------------------------------------------------------- synthetic --

# primes chosen for each digit fold of the MFMA epilogue: 2^31-1 (exact
# shifted sum, Mersenne folds), 2147483629 (general 4-digit Barrett), and for
# the exact 2-digit sum offset by p 2^47: 65521 (largest), 1031 and 1021, and 3
# (the smallest offset above the sum's bound)
@pytest.mark.parametrize("p,b", [(65521, 8), (P31, 16), (P31, 4), (65521, 4), (P31, 8), (65521, 16),
                                 (2147483629, 16), (1031, 8), (1021, 16), (3, 4)])
def test_small_synthetic_plain_and_mfma_vs_oracle(torch_cuda, ctx, p, b):
    n, k, M, Mp = 8, 3, 20, 22
    spec, T, E, N, dr, dc, v, pairs = synthetic_setup(n, k, M, Mp, p, b, tseed=11)
    o = oc.synthetic_op(n, k, M, Mp, p, T, E, dr, dc, 2, pairs.reshape(-1))
    want = o.apply_square(v, b)
    for method in (pg.METHOD_PLAIN, pg.METHOD_MFMA):
        _log(f"small synthetic p={p} b={b} method={method}")
        op = pg.Operator.synthetic(ctx, spec, b, dr, dc, 2, pairs, method=method, tables=(T, E))
        assert op.method == method
        assert (gpu_apply(torch_cuda, op, v) == want).all(), method
        # F (Dc x) without the target preconditioner
        d_in = dev(torch_cuda, v)
        d_out = torch_cuda.empty(op.rows * b, dtype=torch_cuda.int32, device="cuda")
        op.apply_forward(d_in.data_ptr(), d_out.data_ptr(), stream(torch_cuda))
        torch_cuda.cuda.synchronize()
        assert (host(d_out) == o.apply_forward(v, b)).all(), method


@pytest.mark.parametrize("cfg", list(range(MFMA_CONFIGS)))
@pytest.mark.parametrize("p,b,n,k,M,Mp", [(P31, 16, 9, 4, 37, 45), (65521, 8, 10, 3, 70, 29),
                                          (P31, 4, 7, 2, 64, 100), (65521, 16, 8, 3, 50, 90)])
def test_mfma_tile_configs_vs_oracle(torch_cuda, ctx, monkeypatch, cfg, p, b, n, k, M, Mp):
    # every tile configuration the autotuner may choose, on ragged panel sizes
    spec, T, E, N, dr, dc, v, pairs = synthetic_setup(n, k, M, Mp, p, b, tseed=21)
    o = oc.synthetic_op(n, k, M, Mp, p, T, E, dr, dc, 2, pairs.reshape(-1))
    monkeypatch.setenv("PGW_MFMA_CFG", str(cfg))
    op = pg.Operator.synthetic(ctx, spec, b, dr, dc, 2, pairs, method=pg.METHOD_MFMA, tables=(T, E))
    assert (gpu_apply(torch_cuda, op, v) == o.apply_square(v, b)).all()


@pytest.mark.parametrize("cfg", [7, 8, 9])
@pytest.mark.parametrize("p,b,n,k,M,Mp", [(P31, 16, 6, 3, 140, 150), (65521, 8, 6, 2, 136, 290),
                                          (P31, 4, 5, 2, 270, 300)])
def test_mfma_short_last_row_slice_vs_oracle(torch_cuda, ctx, monkeypatch, cfg, p, b, n, k, M, Mp):
    # padded rows that leave a short last row slice, whose workgroups hold
    # one or two row tiles (144 = 128 + 16, 160 = 128 + 32, 304 = 256 + 48,
    # 272 = 256 + 16 rows)
    spec, T, E, N, dr, dc, v, pairs = synthetic_setup(n, k, M, Mp, p, b, tseed=23)
    o = oc.synthetic_op(n, k, M, Mp, p, T, E, dr, dc, 2, pairs.reshape(-1))
    monkeypatch.setenv("PGW_MFMA_CFG", str(cfg))
    op = pg.Operator.synthetic(ctx, spec, b, dr, dc, 2, pairs, method=pg.METHOD_MFMA, tables=(T, E))
    assert (gpu_apply(torch_cuda, op, v) == o.apply_square(v, b)).all()


def test_autotune_keeps_results(torch_cuda, ctx):
    n, k, M, Mp, p, b = 9, 4, 40, 44, P31, 16
    spec, T, E, N, dr, dc, v, pairs = synthetic_setup(n, k, M, Mp, p, b, tseed=2)
    op = pg.Operator.synthetic(ctx, spec, b, dr, dc, 2, pairs, method=pg.METHOD_MFMA, tables=(T, E))
    before = gpu_apply(torch_cuda, op, v)
    choice, ms = op.autotune(2)
    assert len(choice) == 4 and all(0 <= c < MFMA_CONFIGS for c in choice) and all(t > 0 for t in ms)
    assert (gpu_apply(torch_cuda, op, v) == before).all()


@pytest.mark.parametrize("rounds,b,p,n,k,M,Mp", [(2, 16, P31, 8, 3, 20, 22), (1, 8, 65521, 8, 3, 20, 22),
                                                   (2, 3, 1000003, 7, 3, 3, 4), (1, 5, 1009, 7, 3, 3, 4),
                                                   (3, 4, P31, 8, 3, 20, 22), (2, 4, 65521, 9, 4, 30, 31)])
def test_row_mixing_chains_equal_pair_passes(torch_cuda, ctx, monkeypatch, rounds, b, p, n, k, M, Mp):
    # P^T Dr P in one pass along the chains the rounds link (one or two
    # rounds; three keep the pair passes) == the pair passes == the oracle,
    # odd row counts (one row left out of every round) included
    spec, T, E, N, dr, dc, v, pairs = synthetic_setup(n, k, M, Mp, p, b, tseed=7, rounds=rounds)
    o = oc.synthetic_op(n, k, M, Mp, p, T, E, dr, dc, rounds, pairs.reshape(-1))
    want = o.apply_square(v, b)
    _log(f"row-mixing chains rounds={rounds} b={b} p={p} N={N}")
    op = pg.Operator.synthetic(ctx, spec, b, dr, dc, rounds, pairs, tables=(T, E))
    assert (gpu_apply(torch_cuda, op, v) == want).all()  # pair passes
    op.prepare_row_mixing()
    op.prepare_row_mixing()  # a second call changes nothing
    assert (gpu_apply(torch_cuda, op, v) == want).all()  # chains
    monkeypatch.setenv("PGW_MIX_CHAINS", "0")
    op2 = pg.Operator.synthetic(ctx, spec, b, dr, dc, rounds, pairs, tables=(T, E))
    op2.prepare_row_mixing()  # kept on the pair passes
    assert (gpu_apply(torch_cuda, op2, v) == want).all()


def test_odd_rows_and_plain_only_block_widths(torch_cuda, ctx):
    # N odd (one row left out of every round's pairs), b = 3 runs on the plain path
    n, k, M, Mp, p, b = 7, 3, 3, 4, 1000003, 3  # N = 35 * 3 = 105
    spec, T, E, N, dr, dc, v, pairs = synthetic_setup(n, k, M, Mp, p, b, tseed=5)
    assert N % 2 == 1
    o = oc.synthetic_op(n, k, M, Mp, p, T, E, dr, dc, 2, pairs.reshape(-1))
    op = pg.Operator.synthetic(ctx, spec, b, dr, dc, 2, pairs, method=pg.METHOD_AUTO, tables=(T, E))
    assert op.method == pg.METHOD_PLAIN
    assert (gpu_apply(torch_cuda, op, v) == o.apply_square(v, b)).all()
    with pytest.raises(pg.PgwError):
        pg.Operator.synthetic(ctx, spec, b, dr, dc, 2, pairs, method=pg.METHOD_MFMA, tables=(T, E))


def test_c1_full_sequence_mfma_and_plain_vs_oracle(torch_cuda, ctx):
    # n=12, k=3, M=40, M'=48, p=65521, b=8, length 2N/b + 8
    n, k, M, Mp, p, b = 12, 3, 40, 48, 65521, 8
    spec, T, E, N, dr, dc, v, pairs = synthetic_setup(n, k, M, Mp, p, b, tseed=0)
    max_len = 2 * N // b + 8
    o = oc.synthetic_op(n, k, M, Mp, p, T, E, dr, dc, 2, pairs.reshape(-1))
    _log("c1 oracle sequence")
    oseq, ocurv = o.sequence(v, b, max_len)
    for method in (pg.METHOD_MFMA, pg.METHOD_PLAIN):
        _log(f"c1 GPU sequence method={method}")
        op = pg.Operator.synthetic(ctx, spec, b, dr, dc, 2, pairs, method=method, tables=(T, E))
        if method == pg.METHOD_PLAIN:
            op.prepare_row_mixing()  # the one-pass row mixing on this run, the pair passes on the other
        seq, curv = op.sequence(v, max_len)
        assert seq.shape == oseq.shape == (pg.sequence_length(max_len), b * (b + 1) // 2)
        assert (seq == oseq).all(), method
        assert (curv == ocurv).all(), method


def test_gram_upper_vs_oracle(torch_cuda, ctx):
    for p, b, rows in [(P31, 16, 100003), (65521, 8, 5000), (109, 3, 777)]:
        n, k, M, Mp = 6, 3, rows, 1  # only the shape (cols = C(6,3)*rows) matters here
        spec = pg.SyntheticSpec(n, k, M, Mp, p, 0, 1)
        N = comb(n, k) * M
        dr = np.ones(N, np.uint32)
        op = pg.Operator.synthetic(ctx, spec, b, dr, dr, 0, None, method=pg.METHOD_PLAIN)
        rng = np.random.default_rng(b)
        x = rng.integers(0, p, (N, b)).astype(np.uint32)
        y = rng.integers(0, p, (N, b)).astype(np.uint32)
        d_x, d_y = dev(torch_cuda, x), dev(torch_cuda, y)
        d_o = torch_cuda.zeros(b * (b + 1) // 2, dtype=torch_cuda.int32, device="cuda")
        op.gram_upper(d_x.data_ptr(), d_y.data_ptr(), d_o.data_ptr(), stream(torch_cuda))
        torch_cuda.cuda.synchronize()
        assert (host(d_o) == oc.gram_upper(x, y, N, b, p)).all(), (p, b)


def test_c2_single_sector_apply_mfma_equals_plain_and_oracle(torch_cuda, ctx):
    # n=20, k=5, M=200, M'=220 (N = 3,100,800), p = 2^31-1, b = 16, V from seed 2
    n, k, M, Mp, p, b = 20, 5, 200, 220, P31, 16
    spec, T, E, N, dr, dc, v, pairs = synthetic_setup(n, k, M, Mp, p, b, tseed=0, seed=2)
    assert N == 3100800
    torch = torch_cuda
    ops = {m: pg.Operator.synthetic(ctx, spec, b, dr, dc, 2, pairs, method=m, tables=(T, E))
           for m in (pg.METHOD_MFMA, pg.METHOD_PLAIN)}
    ops[pg.METHOD_MFMA].prepare_row_mixing()  # one-pass row mixing against the plain path's pair passes
    rng = np.random.default_rng(2)
    y = rng.integers(0, p, (N, b)).astype(np.uint32)
    outs = {}
    for m, op in ops.items():
        _log(f"c2 apply method={m}")
        outs[m] = (gpu_apply(torch, op, v), gpu_apply(torch, op, y))
    assert (outs[pg.METHOD_MFMA][0] == outs[pg.METHOD_PLAIN][0]).all()
    assert (outs[pg.METHOD_MFMA][1] == outs[pg.METHOD_PLAIN][1]).all()
    sv, sy = outs[pg.METHOD_MFMA]
    # S symmetric: <Sv, y> == <v, Sy> mod p
    def dot(a, c):  # terms reduced below 2^31, 50M of them stay below 2^63
        return int(((a.astype(np.int64) * c.astype(np.int64)) % p).sum() % p)

    assert dot(sv, y.reshape(-1)) == dot(v, sy)
    # linearity: S(v + y) = Sv + Sy
    vy = ((v.astype(np.uint64) + y.reshape(-1)) % p).astype(np.uint32)
    svy = gpu_apply(torch, ops[pg.METHOD_MFMA], vy)
    assert (svy == ((sv.astype(np.uint64) + sy) % p).astype(np.uint32)).all()
    _log("c2 oracle apply")
    o = oc.synthetic_op(n, k, M, Mp, p, T, E, dr, dc, 2, pairs.reshape(-1))
    assert (o.apply_square(v, b) == sv).all()


def test_c4_large_sector_mfma_equals_plain_symmetry_linearity(torch_cuda, ctx):
    # n=24, k=6, M=400, M'=440 (N = 53,838,400), p = 2^31-1, b = 16, tables from seed 4:
    # full-size checks through size-independent properties (the oracle would need minutes)
    n, k, M, Mp, p, b = 24, 6, 400, 440, P31, 16
    _log("c4 setup")
    spec, T, E, N, dr, dc, v, pairs = synthetic_setup(n, k, M, Mp, p, b, tseed=4, seed=4)
    assert N == 53838400
    torch = torch_cuda
    rng = np.random.default_rng(4)
    y = rng.integers(0, p, N * b, dtype=np.uint32)
    mf = pg.Operator.synthetic(ctx, spec, b, dr, dc, 2, pairs, method=pg.METHOD_MFMA, tables=(T, E))
    _log("c4 MFMA applies")
    sv = gpu_apply(torch, mf, v)
    sy = gpu_apply(torch, mf, y)
    vy = ((v.astype(np.uint64) + y) % p).astype(np.uint32)
    svy = gpu_apply(torch, mf, vy)
    mf.close()
    del vy
    assert (svy == ((sv.astype(np.uint64) + sy) % p).astype(np.uint32)).all()
    del svy

    def dot(a, c):  # 861M terms below 2^31 each: the int64 sum stays below 2^61
        acc = 0
        for s0 in range(0, a.size, 1 << 26):
            acc += int(((a[s0:s0 + (1 << 26)].astype(np.int64) *
                         c[s0:s0 + (1 << 26)].astype(np.int64)) % p).sum())
        return acc % p

    assert dot(sv, y) == dot(v, sy)
    _log("c4 plain apply")
    pl = pg.Operator.synthetic(ctx, spec, b, dr, dc, 2, pairs, method=pg.METHOD_PLAIN, tables=(T, E))
    assert (gpu_apply(torch, pl, v) == sv).all()
    pl.close()

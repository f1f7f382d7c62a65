"""GPU: the block projection on int8 MFMA (pgw_operator_set_gram(MFMA)) is
bit-identical to the vector-ALU projection and to the oracle's
orc_gram_upper, for every block width the MFMA path takes (b = 16 uses the
whole 16 x 16 product, b = 8 and 4 its diagonal blocks), for one- and
two-limb primes, ragged row counts, and blocks whose digits make every
accumulator grow as fast as it can (the periodic folds must keep them exact)."""
import os
import sys
from math import comb

import numpy as np
import pytest

import oracle_ctypes as oc
import prym_green_ancillary_amd as pg

pytestmark = pytest.mark.gpu

P31 = 2147483647


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


def shape_op(ctx, p, b, rows):
    # only the operator's column count (C(6,3) * rows) and b, p matter to a projection
    spec = pg.SyntheticSpec(6, 3, rows, 1, p, 0, 1)
    N = comb(6, 3) * rows
    ones = np.ones(N, np.uint32)
    return pg.Operator.synthetic(ctx, spec, b, ones, ones, 0, None, method=pg.METHOD_PLAIN), N


def project(torch, op, x, y, method):
    b = x.shape[1]
    d_x = torch.from_numpy(x.view(np.int32)).cuda()
    d_y = torch.from_numpy(y.view(np.int32)).cuda()
    d_o = torch.zeros(b * (b + 1) // 2, dtype=torch.int32, device="cuda")
    op.set_gram(method)
    op.gram_upper(d_x.data_ptr(), d_y.data_ptr(), d_o.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return d_o.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("p,b,rows", [
    (P31, 16, 100003), (2147483629, 16, 7777), (65521, 8, 5000), (65521, 4, 12345),
    (P31, 8, 999), (P31, 4, 31), (661, 4, 5), (109, 16, 1), (1009, 8, 43210)])
def test_gram_mfma_equals_valu_and_oracle(torch_cuda, ctx, p, b, rows):
    op, N = shape_op(ctx, p, b, rows)
    rng = np.random.default_rng(rows + b)
    x = rng.integers(0, p, (N, b)).astype(np.uint32)
    y = rng.integers(0, p, (N, b)).astype(np.uint32)
    _log(f"gram p={p} b={b} N={N}")
    want = oc.gram_upper(x, y, N, b, p)
    assert (project(torch_cuda, op, x, y, pg.METHOD_PLAIN) == want).all()
    assert (project(torch_cuda, op, x, y, pg.METHOD_MFMA) == want).all()
    op.close()


@pytest.mark.parametrize("p,b", [(P31, 16), (2147483629, 8), (65521, 4)])
def test_gram_mfma_folds_keep_long_sums_exact(torch_cuda, ctx, p, b):
    # one workgroup (PGW_GRAM_MAX_GRID=1) walks 2^20 rows: thousands of chunks
    # per wave, so the shift accumulators are folded many times; every entry
    # is the largest representable digit pattern (all digits 127), so each
    # accumulator grows by the most a chunk can add
    nl = 2 if p <= 65536 else 4
    top = 127 * (256 ** nl - 1) // 255  # digits (127, ..., 127), a residue below p
    assert top < p
    rows = (1 << 20) // comb(6, 3) + 1
    op, N = shape_op(ctx, p, b, rows)
    x = np.full((N, b), top, np.uint32)
    rng = np.random.default_rng(b)
    y = x.copy()
    y[rng.integers(0, N, 64), rng.integers(0, b, 64)] = rng.integers(0, p, 64).astype(np.uint32)
    _log(f"gram folds p={p} b={b} N={N}")
    want = oc.gram_upper(x, y, N, b, p)
    os.environ["PGW_GRAM_MAX_GRID"] = "1"
    try:
        got = project(torch_cuda, op, x, y, pg.METHOD_MFMA)
    finally:
        del os.environ["PGW_GRAM_MAX_GRID"]
    assert (got == want).all()
    op.close()


def test_gram_method_choice_is_checked(ctx):
    op, _ = shape_op(ctx, 109, 3, 4)
    with pytest.raises(pg.PgwError, match="b in"):
        op.set_gram(pg.METHOD_MFMA)
    op.set_gram(pg.METHOD_PLAIN)
    op.set_gram(pg.METHOD_AUTO)
    op.close()

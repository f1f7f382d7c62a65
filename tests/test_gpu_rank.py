"""End to end on the GPU: sequence on the device (HIP library), rank and
generator on the host (pgw_rank_*), kernel vectors back on the device.

Checks: the genus-12 sector ranks of g12-reference.json, the kernel-vector
certificate F (Dc D) = 0 with rank D = the sector's nullity, and for the small
synthetic sector (n=12, k=3, M=40, M'=48, p=65521, b=8, length 2N/b + 8) a rank
lower bound equal to the upper bound dim Lambda^2 (x) R_(d+1) = 3168, which
certifies the rank of the Gram operator.
"""
import sys
from math import comb

import numpy as np
import pytest

import oracle_ctypes as oc
import prym_green_ancillary_amd as pg

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    c = pg.Context(0)
    yield c
    c.close()


def test_g12_gpu_sequence_rank_and_kernel_vectors(ctx, g12, g12_reference):
    p, b = g12["modulus"], 4
    for prof in g12_reference["sector_profiles"]:
        sector = prof["sector"]
        _log(f"g12 sector {sector}: GPU sequence, host rank, GPU kernel vectors")
        rows = prof["rows"]
        cols = prof["columns"]
        dr, dc, v = pg.reference_random_state(1, sector, p, rows, cols, b)
        pairs = pg.rowmix_plan(pg.rowmix_realized_seed(10001, "eliminated", sector, rows, p), rows, 2, p)
        op = pg.Operator.cyclic(ctx, g12, sector, b, dr, dc, 2, pairs)
        assert (op.rows, op.cols) == (rows, cols)
        seq, _ = op.sequence(v, 2 * cols // b + 50)
        r = pg.RankComputation(p, b, rows, cols)
        st = r.step(seq)
        assert st.rank_lower_bound == prof["rank"] and st.status == pg.RANK_SUCCESS, prof
        gen = r.generator()
        _, rank_o, spread_o, gen_o = oc.bcw_rank(p, b, seq)
        assert (rank_o, spread_o) == (st.rank_lower_bound, st.delta_spread)
        assert np.array_equal(gen, gen_o)
        d, prod = op.kernel_vectors(v, gen.reshape(gen.shape[0], b * b))
        assert not prod.any(), prof
        assert oc.rank_mod_p(d.reshape(cols, b), p) == prof["nullity"], prof
        op.close()


@pytest.mark.parametrize("method", [pg.METHOD_MFMA, pg.METHOD_PLAIN])
def test_small_synthetic_sector_rank_certified(ctx, method):
    n, k, M, Mp, p, b = 12, 3, 40, 48, 65521, 8
    spec = pg.SyntheticSpec(n, k, M, Mp, p, 0, 1)
    T, E = spec.tables()
    N = comb(n, k) * M
    dr, dc, v = pg.reference_random_state(1, 0, p, N, N, b)
    pairs = pg.rowmix_plan(pg.rowmix_realized_seed(10001, "synthetic", 0, N, p), N, 2, p)
    _log(f"synthetic N={N} b={b} method={method}: GPU sequence")
    op = pg.Operator.synthetic(ctx, spec, b, dr, dc, 2, pairs, method=method, tables=(T, E))
    seq, _ = op.sequence(v, 2 * N // b + 8)
    op.close()
    _log("host rank")
    r = pg.RankComputation(p, b, N, N)
    st = r.step(seq)
    upper = comb(n, k - 1) * Mp
    assert st.rank_lower_bound == upper == 3168
    assert st.status == pg.RANK_SUCCESS and st.delta_spread >= 10

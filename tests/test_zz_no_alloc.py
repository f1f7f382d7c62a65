"""GPU: the allocation-free Gram entry point (pgw_gram_upper_noalloc) against
the operator's pgw_gram_upper and the oracle, on two calls in a row, the
second with fewer rows (the fixed scratch keeps the first call's partials)."""
import numpy as np
import pytest

import oracle_ctypes as oc
import prym_green_ancillary_amd as pg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("p,b,shapes", [
    (2147483647, 16, [(9, 3, 40), (7, 3, 12)]),
    (65521, 5, [(8, 3, 30), (6, 2, 7)]),
    (109, 8, [(10, 4, 25), (5, 2, 3)]),
])
def test_noalloc_gram_matches_operator_gram(p, b, shapes):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    ctx = pg.Context(0)
    rng = np.random.default_rng(p % 1000 + b)
    for n, k, M in shapes:  # the second call has fewer rows than the first
        spec = pg.SyntheticSpec(n, k, M, M + 2, p, 5, 6)
        N = spec.cols
        print(f"noalloc gram p={p} b={b} N={N}", flush=True)
        dr = rng.integers(1, p, N).astype(np.uint32)
        dc = rng.integers(1, p, N).astype(np.uint32)
        op = pg.Operator.synthetic(ctx, spec, b, dr, dc, 0, None, method=pg.METHOD_PLAIN)
        x = rng.integers(0, p, N * b).astype(np.uint32)
        y = rng.integers(0, p, N * b).astype(np.uint32)
        dx = torch.from_numpy(x.view(np.int32)).cuda()
        dy = torch.from_numpy(y.view(np.int32)).cuda()
        ut = b * (b + 1) // 2
        want = torch.zeros(ut, dtype=torch.int32, device="cuda")
        got = torch.full((ut,), -1, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        op.gram_upper(dx.data_ptr(), dy.data_ptr(), want.data_ptr())
        ctx.synchronize()
        ctx.gram_upper_noalloc(p, b, N, dx.data_ptr(), dy.data_ptr(), got.data_ptr())
        ctx.synchronize()
        w = want.cpu().numpy().view(np.uint32)
        g = got.cpu().numpy().view(np.uint32)
        assert np.array_equal(g, w), (p, b, N)
        assert np.array_equal(g, oc.gram_upper(x, y, N, b, p)), (p, b, N)
        op.close()
    with pytest.raises(pg.PgwError, match="block width"):
        ctx.gram_upper_noalloc(p, 17, 4, dx.data_ptr(), dy.data_ptr(), got.data_ptr())
    ctx.close()

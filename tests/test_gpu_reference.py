"""GPU: the eliminated forward F (Dc x) of the product (pgw_apply_forward,
through the C ABI) on the reference's deterministic blocks equals what the
reference's own CPU code computes (golden vectors from
cpu_eliminated_forward, cuda/cuprym_cyclic.cu:1775-1799, written by
tests/golden/make_ref_golden.sh), for every genus-12 sector and three
genus-20 sectors."""
import json

import numpy as np
import pytest

import prym_green_ancillary_amd as pg
from test_reference_golden import GOLDEN, deterministic_block, fnv_values, reference_cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("genus", [12, 20])
def test_gpu_eliminated_forward_matches_reference_cpu(genus):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    inst = json.loads((GOLDEN / f"g{genus}.json").read_text())
    ctx = pg.Context(0)
    for rec in reference_cases(genus):
        s, p, b = rec["sector"], rec["prime"], rec["b"]
        print(f"reference forward g{genus} sector {s}", flush=True)
        dr, dc, _ = pg.reference_random_state(1, s, p, rec["rows"], rec["cols"], b)
        pairs = pg.rowmix_plan(pg.rowmix_realized_seed(10001, "eliminated", s, rec["rows"], p),
                               rec["rows"], 2, p)
        op = pg.Operator.cyclic(ctx, inst, s, b, dr, dc, 2, pairs)
        assert (op.rows, op.cols) == (rec["rows"], rec["cols"])
        for app in rec["applies"]:
            x = deterministic_block(op.cols, b, p, app["salt"])
            dx = torch.from_numpy(x.view(np.int32).copy()).cuda()
            dy = torch.zeros(op.rows * b, dtype=torch.int32, device="cuda")
            torch.cuda.synchronize()
            op.apply_forward(dx.data_ptr(), dy.data_ptr())
            ctx.synchronize()
            f = dy.cpu().numpy().view(np.uint32)
            assert f[:16].tolist() == app["forward_head"] and fnv_values(f) == app["forward_fnv1a64"], \
                (genus, s, app["salt"])
        op.close()
    ctx.close()

"""The product's exact host apply of the cyclic operators (pgw_cyclic_apply_cpu,
the counterpart of cpu_apply_eliminated_square, cuda/cuprym_cyclic.cu:
1591-1966) and its deterministic blocks (deterministic_dense_block,
:1410-1420), which the CLI's --validate-rowmix compares with the device apply.

* the deterministic blocks reproduce the reference's own input hashes
  (ref_forward_g12/g20.json, written by the reference's code) and the
  splitmix64 restatement in test_reference_golden;
* the host apply equals the oracle's apply_square on every genus-12 sector,
  on genus-20 sector 0, and on the augmented and replacement variants;
* S is symmetric (<S u, w> = <u, S w>), and the apply refuses bad inputs.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle_ctypes as oc
import prym_green_ancillary_amd as pg
from test_deformation import augmented_block, replacement_parts
from test_reference_golden import deterministic_block, fnv_values

GOLDEN = Path(__file__).resolve().parent / "golden"


def _paper_state(inst, sector, op_name, rows, cols, b, rounds=2):
    p = inst["modulus"]
    dr, dc, _ = pg.reference_random_state(1, sector, p, rows, cols, b)
    pairs = pg.rowmix_plan(pg.rowmix_realized_seed(10001, op_name, sector, rows, p), rows, rounds, p)
    return dr, dc, pairs


@pytest.mark.parametrize("genus", [12, 20])
def test_deterministic_block_matches_reference_hashes(genus):
    rec = json.loads((GOLDEN / f"ref_forward_g{genus}.json").read_text())["sectors"][0]
    cols, b, p = rec["cols"], rec["b"], rec["prime"]
    for app in rec["applies"]:
        x = pg.deterministic_block(cols, b, p, app["salt"])
        assert fnv_values(x) == app["input_fnv1a64"]
    for salt in (0x404, 0x505, 0x606, 0xA11CE001):
        assert (pg.deterministic_block(37, 3, 661, salt) == deterministic_block(37, 3, 661, salt)).all()


@pytest.mark.parametrize("width", [4, 3])
def test_host_apply_equals_oracle_on_every_g12_sector(g12, width):
    p = g12["modulus"]
    for sector in range(g12["cyclic"]["order"]):
        o = oc.cyclic_op(g12, sector)
        d, keep = pg.Operator.cyclic_desc(g12, sector, width)
        assert pg.Operator.dims(d) == (o.rows, o.cols)
        dr, dc, pairs = _paper_state(g12, sector, "eliminated", o.rows, o.cols, width)
        o.set_preconditioners(dr, dc, 2, pairs.reshape(-1))
        for salt in (0x404, 0x505, 0x606):
            x = pg.deterministic_block(o.cols, width, p, salt)
            got = pg.cyclic_apply_cpu(d, dr, dc, 2, pairs, x, width)
            assert (got == o.apply_square(x, width)).all(), (sector, salt)


def test_host_apply_equals_oracle_without_row_mixing(g12):
    p, b, sector = g12["modulus"], 2, 3
    o = oc.cyclic_op(g12, sector)
    d, keep = pg.Operator.cyclic_desc(g12, sector, b)
    dr, dc, _ = pg.reference_random_state(5, sector, p, o.rows, o.cols, b)
    o.set_preconditioners(dr, dc, 0, None)
    x = pg.deterministic_block(o.cols, b, p, 0x77)
    assert (pg.cyclic_apply_cpu(d, dr, dc, 0, None, x, b) == o.apply_square(x, b)).all()


def test_host_apply_equals_oracle_on_g20_sector0():
    inst = json.loads((GOLDEN / "g20.json").read_text())
    p, b, sector = inst["modulus"], 4, 0
    o = oc.cyclic_op(inst, sector)
    d, keep = pg.Operator.cyclic_desc(inst, sector, b)
    dr, dc, pairs = _paper_state(inst, sector, "eliminated", o.rows, o.cols, b)
    o.set_preconditioners(dr, dc, 2, pairs.reshape(-1))
    x = pg.deterministic_block(o.cols, b, p, 0x404)
    assert (pg.cyclic_apply_cpu(d, dr, dc, 2, pairs, x, b) == o.apply_square(x, b)).all()


@pytest.mark.parametrize("variant,sector", [("augmented", 1), ("augmented", 5), ("replacement", 0)])
def test_host_apply_equals_oracle_on_deformation_variants(g12, g12_reference, variant, sector):
    p, b = g12["modulus"], 3
    if variant == "augmented":
        dense, _ = augmented_block(g12, g12_reference, sector)
        drops = (0, 0)
    else:
        dense, J = replacement_parts(g12_reference)
        drops = tuple(J)
    o = oc.cyclic_op(g12, sector).deform(variant, dense, drops)
    d, keep = pg.Operator.deformation_desc(g12, sector, b, variant, dense, drops)
    assert pg.Operator.dims(d) == (o.rows, o.cols)
    dr, dc, pairs = _paper_state(g12, sector, variant, o.rows, o.cols, b)
    o.set_preconditioners(dr, dc, 2, pairs.reshape(-1))
    for salt in (0x404, 0x606):
        x = pg.deterministic_block(o.cols, b, p, salt)
        assert (pg.cyclic_apply_cpu(d, dr, dc, 2, pairs, x, b) == o.apply_square(x, b)).all()


def test_host_apply_is_symmetric(g12):
    p, b, sector = g12["modulus"], 4, 2
    d, keep = pg.Operator.cyclic_desc(g12, sector, b)
    rows, cols = pg.Operator.dims(d)
    dr, dc, pairs = _paper_state(g12, sector, "eliminated", rows, cols, b)
    u = pg.deterministic_block(cols, b, p, 1)
    w = pg.deterministic_block(cols, b, p, 2)
    su = pg.cyclic_apply_cpu(d, dr, dc, 2, pairs, u, b).astype(np.int64)
    sw = pg.cyclic_apply_cpu(d, dr, dc, 2, pairs, w, b).astype(np.int64)
    assert ((su * w.astype(np.int64)).sum() - (u.astype(np.int64) * sw).sum()) % p == 0


def test_host_apply_refuses_bad_inputs(g12):
    p, b = g12["modulus"], 4
    d, keep = pg.Operator.cyclic_desc(g12, 0, b)
    rows, cols = pg.Operator.dims(d)
    dr, dc, pairs = _paper_state(g12, 0, "eliminated", rows, cols, b)
    x = pg.deterministic_block(cols, b, p, 9)
    with pytest.raises(pg.PgwError, match="block width"):
        pg.cyclic_apply_cpu(d, dr, dc, 2, pairs, x, 0)
    bad = x.copy()
    bad[5] = p
    with pytest.raises(pg.PgwError, match="reduced modulo p"):
        pg.cyclic_apply_cpu(d, dr, dc, 2, pairs, bad, b)
    zero_dr = dr.copy()
    zero_dr[0] = 0
    with pytest.raises(pg.PgwError, match="row preconditioner"):
        pg.cyclic_apply_cpu(d, zero_dr, dc, 2, pairs, x, b)
    clash = pairs.copy()
    clash[1, 0] = clash[0, 0]
    with pytest.raises(pg.PgwError, match="disjoint"):
        pg.cyclic_apply_cpu(d, dr, dc, 2, clash, x, b)
    spec = pg.SyntheticSpec(6, 3, 4, 5, 65521, 7, 8)
    T, E = spec.tables()
    sd = pg.Operator.synthetic_desc(spec, 2, pg.METHOD_PLAIN, T, E)
    with pytest.raises(pg.PgwError, match="cyclic operators"):
        pg.cyclic_apply_cpu(sd, dr, dc, 0, None, x, 2)

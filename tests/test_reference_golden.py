"""Golden vectors written by the reference's own host C++ (oracle/ref_build.sh
compiles cuda/cuprym_cyclic.cu's host-only ranges; tests/golden/
make_ref_golden.sh runs them).  They pin, for both the product's host recipe
and the oracle restatement:

* seeded_sector_value, rowmix_realized_seed, generate_rowmix_plan (pairs
  sampled at both ends of every round) and rowmix_plan_hash at the genus-12
  and genus-20 sector heights for p = 109, 661, 65521 and the three operator
  names (cuda/cuprym_cyclic.cu:1143-1287);
* the mt19937_64 Dr / Dc / V draws of run_eliminated_sector
  (cuda/cuprym_cyclic.cu:2196-2199, cuda/matrices.h:44-56);
* the eliminated forward F (Dc x) and P^T Dr P F (Dc x) of every genus-12
  sector and three genus-20 sectors on the reference's deterministic blocks
  (cpu_eliminated_forward / cpu_apply_target_preconditioner,
  cuda/cuprym_cyclic.cu:1685-1799), against the oracle here and against the
  GPU in tests/test_gpu_reference.py.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle_ctypes as oc
import prym_green_ancillary_amd as pg

GOLDEN = Path(__file__).resolve().parent / "golden"
MASK = (1 << 64) - 1


def _load(name):
    return json.loads((GOLDEN / name).read_text())


def test_seeded_sector_value_matches_reference():
    for seed, sector, want in _load("ref_plan.json")["seeded_sector_value"]:
        assert pg.lib().pgw_seeded_sector_value(seed, sector) == want
        assert oc.lib().orc_seeded_sector_value(seed, sector) == want


def _samples(pairs, rows):
    pc = rows // 2
    pr = np.asarray(pairs, dtype=np.uint64).reshape(2, pc, 4)
    head = pr[:, :3, :].reshape(2, -1)
    return (np.concatenate([head[0], head[1]]).tolist(),
            np.concatenate([pr[0, -1], pr[1, -1]]).tolist())


@pytest.mark.parametrize("rec", _load("ref_plan.json")["plans"],
                         ids=lambda r: f"{r['op']}-{r['rows']}-{r['prime']}-{r['sector']}")
def test_rowmix_plan_and_hash_match_reference(rec):
    op, rows, p, sector = rec["op"], rec["rows"], rec["prime"], rec["sector"]
    # product host recipe
    realized = pg.rowmix_realized_seed(rec["seed"], op, sector, rows, p)
    assert realized == rec["realized_seed"]
    pairs = pg.rowmix_plan(realized, rows, rec["rounds"], p)
    head, tail = _samples(pairs, rows)
    assert head == rec["first_pairs"] and tail == rec["last_pairs"]
    assert pg.rowmix_plan_hash(rows, rec["rounds"], rec["seed"], realized, pairs) == rec["plan_hash"]
    # oracle restatement
    o_realized, o_pairs, o_hash = oc.rowmix_plan(rec["seed"], op, sector, rows, p, rec["rounds"])
    assert o_realized == rec["realized_seed"] and o_hash == rec["plan_hash"]
    o_head, o_tail = _samples(o_pairs, rows)
    assert o_head == rec["first_pairs"] and o_tail == rec["last_pairs"]


def test_reference_random_state_matches_reference():
    for rec in _load("ref_state.json")["states"]:
        args = (rec["seed"], rec["sector"], rec["prime"], rec["rows"], rec["cols"], rec["b"])
        for dr, dc, v in (pg.reference_random_state(*args), oc.reference_random_state(*args)):
            assert dr.tolist() == rec["dr"] and dc.tolist() == rec["dc"] and v.tolist() == rec["v"]


def deterministic_block(rows, cols, p, salt):
    """deterministic_dense_block (cuda/cuprym_cyclic.cu:1410-1420): splitmix64
    from salt ^ (0xD6E8FEB86659FD93 + rows*1315423911 + cols*2654435761)."""
    state = (salt ^ ((0xD6E8FEB86659FD93 + rows * 1315423911 + cols * 2654435761) & MASK)) & MASK
    out = np.empty(rows * cols, np.uint32)
    for q in range(rows * cols):
        state = (state + 0x9E3779B97F4A7C15) & MASK
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        out[q] = (z ^ (z >> 31)) % p
    return out


def fnv_values(v) -> int:
    return oc.lib().orc_fnv1a64_u32_vector(np.ascontiguousarray(v, dtype=np.uint32), np.size(v))


def target_precond(t, pairs, rounds, dr, p, b):
    """P^T Dr P on a rows x b block (cyclic_rowmix_pair_kernel rounds, then Dr,
    then the transposed rounds in reverse: cuda/cuprym_cyclic.cu:1687-1702)."""
    x = t.reshape(-1, b).astype(object)
    rows = x.shape[0]
    pr = np.asarray(pairs).reshape(rounds, rows // 2, 4).astype(object)
    for r in range(rounds):
        f, s, a, bb = pr[r, :, 0], pr[r, :, 1], pr[r, :, 2], pr[r, :, 3]
        xf, xs = x[f.astype(int)].copy(), x[s.astype(int)].copy()
        x[f.astype(int)] = (xf + a[:, None] * xs) % p
        x[s.astype(int)] = (bb[:, None] * xf + ((1 + a * bb) % p)[:, None] * xs) % p
    x = x * dr.astype(object)[:, None] % p
    for r in reversed(range(rounds)):
        f, s, a, bb = pr[r, :, 0], pr[r, :, 1], pr[r, :, 2], pr[r, :, 3]
        xf, xs = x[f.astype(int)].copy(), x[s.astype(int)].copy()
        x[f.astype(int)] = (xf + bb[:, None] * xs) % p
        x[s.astype(int)] = (a[:, None] * xf + ((1 + a * bb) % p)[:, None] * xs) % p
    return np.asarray(x % p, dtype=np.uint32).reshape(-1)


def reference_cases(genus):
    return _load(f"ref_forward_g{genus}.json")["sectors"]


@pytest.mark.parametrize("genus", [12, 20])
def test_oracle_eliminated_forward_matches_reference_cpu(genus):
    inst = json.loads((GOLDEN / f"g{genus}.json").read_text())
    for rec in reference_cases(genus):
        s, p, b = rec["sector"], rec["prime"], rec["b"]
        o = oc.cyclic_op(inst, s)
        assert (o.rows, o.cols) == (rec["rows"], rec["cols"])
        dr, dc, _ = oc.reference_random_state(1, s, p, o.rows, o.cols, b)
        realized, pairs, plan_hash = oc.rowmix_plan(10001, "eliminated", s, o.rows, p, 2)
        assert plan_hash == rec["plan_hash"]
        o.set_preconditioners(dr, dc, 2, pairs)
        for app in rec["applies"]:
            x = deterministic_block(o.cols, b, p, app["salt"])
            assert fnv_values(x) == app["input_fnv1a64"]
            f = o.apply_forward(x, b)
            assert f[:16].tolist() == app["forward_head"] and fnv_values(f) == app["forward_fnv1a64"]
            mixed = target_precond(f, pairs, 2, dr, p, b)
            assert mixed[:16].tolist() == app["mixed_head"] and fnv_values(mixed) == app["mixed_fnv1a64"]


def test_oracle_sweep_partial_rows_equal_the_dense_sweeps():
    # orc_sweep_partial (the sampled CPU baseline) over row ranges equals the
    # dense exterior-algebra matrices D, E, E^T, D^T applied to the whole block
    from math import comb
    from test_oracle import _dense_synthetic
    n, k, M, Mp, p, b = 6, 3, 4, 5, 65521, 3
    T = oc.synthetic_table(7, p, n * Mp * M)
    E = oc.synthetic_table(8, p, n * Mp * M)
    N, R = comb(n, k) * M, comb(n, k - 1) * Mp
    dr, dc, x = oc.reference_random_state(1, 0, p, N, N, b)
    y = oc.reference_random_state(2, 0, p, N, R, b)[2]
    o = oc.synthetic_op(n, k, M, Mp, p, T, E, dr, dc)
    D, Em = _dense_synthetic(n, k, M, Mp, T, E, p)
    dense = {0: (D, x), 1: (Em, y), 2: (Em.T, x), 3: (D.T, y)}
    for which, (A, xin) in dense.items():
        rows = oc.lib().orc_sweep_rows(o.h, which)
        assert rows == A.shape[0]
        parts = np.zeros(rows * b, np.uint32)
        for r0, r1 in ((0, 7), (7, rows // 2), (rows // 2, rows)):
            oc.lib().orc_sweep_partial(o.h, which, xin, parts, b, r0, r1)
        want = A.dot(xin.reshape(-1, b).astype(object)) % p
        assert (parts.reshape(-1, b).astype(object) == want).all(), which

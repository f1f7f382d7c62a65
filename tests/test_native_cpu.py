"""CPU-side checks of libpgw.so: it loads, exports the C ABI of include/pgw.h,
and its host recipe (RNG streams, row-mix plan, hashes, operator shapes) is
bit-identical to the oracle's.  No GPU call is made here."""
import re
from math import comb

import numpy as np
import pytest

import oracle_ctypes as oc
import prym_green_ancillary_amd as pg

HEADER = oc.ROOT / "include" / "pgw.h"


def test_library_exports_every_declared_symbol():
    declared = set(re.findall(r"\b(pgw_[a-z0-9_]+)\s*\(", HEADER.read_text()))
    declared.discard("pgw_progress_fn")
    assert declared, "no declarations found"
    L = pg.lib()
    for name in sorted(declared):
        assert hasattr(L, name), name
    assert set(pg.EXPORTED) == declared


def test_one_hip_runtime_per_process():
    # libpgw shares the HIP and HSA runtimes torch brought in: a process
    # holds one copy of each, whichever of the two was asked for first
    import subprocess
    import sys
    code = ("import prym_green_ancillary_amd as pg; pg.lib(); import torch; "
            "m = open('/proc/self/maps').read().splitlines(); "
            "print(len({l.split()[-1] for l in m if 'libamdhip64' in l}), "
            "len({l.split()[-1] for l in m if 'libhsa-runtime64' in l}))")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                         cwd=str(HEADER.parents[1]))
    assert out.returncode == 0, out.stderr[-2000:]
    assert out.stdout.split() == ["1", "1"]


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(pg, "LIB_PATH", tmp_path / "absent.so")
    monkeypatch.setattr(pg, "_lib", None)
    with pytest.raises(pg.NativeLibraryError):
        pg.lib()


@pytest.mark.parametrize("seed,sector,p,rows,cols,b", [(1, 0, 109, 106, 84, 4), (1, 5, 65521, 50, 40, 8),
                                                        (2, 0, 2147483647, 33, 31, 16), (7, 3, 661, 10, 12, 1)])
def test_reference_random_state_matches_oracle(seed, sector, p, rows, cols, b):
    got = pg.reference_random_state(seed, sector, p, rows, cols, b)
    want = oc.reference_random_state(seed, sector, p, rows, cols, b)
    for g, w in zip(got, want):
        assert (g == w).all()


@pytest.mark.parametrize("rows,p,rounds", [(106, 109, 2), (101, 65521, 2), (64, 2147483647, 3), (3, 97, 1)])
def test_rowmix_plan_and_hash_match_oracle(rows, p, rounds):
    realized = pg.rowmix_realized_seed(10001, "eliminated", 4, rows, p)
    o_realized, o_pairs, o_hash = oc.rowmix_plan(10001, "eliminated", 4, rows, p, rounds)
    assert realized == o_realized
    pairs = pg.rowmix_plan(realized, rows, rounds, p)
    assert (pairs.reshape(-1) == o_pairs).all()
    assert pg.rowmix_plan_hash(rows, rounds, 10001, realized, pairs) == o_hash


def test_fnv_and_tables_match_oracle():
    v = np.arange(1000, dtype=np.uint32) * 7919
    assert pg.fnv1a64_u32(v) == oc.lib().orc_fnv1a64_u32_vector(v, v.size)
    for seed, p in [(0, 65521), (1, 2147483647), (4, 109)]:
        assert (pg.synthetic_table(seed, p, 5000) == oc.synthetic_table(seed, p, 5000)).all()


def test_operator_dims_host_only(g12):
    for sector in range(6):
        d, keep = pg.Operator.cyclic_desc(g12, sector, 4)
        rows, cols = pg.Operator.dims(d)
        op = oc.cyclic_op(g12, sector)
        assert (rows, cols) == (op.rows, op.cols)
    spec = pg.SyntheticSpec(12, 3, 40, 48, 65521, 0, 1)
    td, te = spec.tables()
    d = pg.Operator.synthetic_desc(spec, 8, pg.METHOD_AUTO, td, te)
    assert pg.Operator.dims(d) == (comb(12, 3) * 40, comb(12, 3) * 40)


def test_invalid_descriptions_are_rejected(g12):
    d, keep = pg.Operator.cyclic_desc(g12, 0, 4)
    d.prime = 110  # not prime
    with pytest.raises(pg.PgwError):
        pg.Operator.dims(d)
    d, keep = pg.Operator.cyclic_desc(g12, 6, 4)  # sector outside 0..r-1
    with pytest.raises(pg.PgwError, match="sector"):
        pg.Operator.dims(d)
    d, keep = pg.Operator.cyclic_desc(g12, 0, 17)
    with pytest.raises(pg.PgwError, match="block"):
        pg.Operator.dims(d)

"""Deformation variants of the eliminated cyclic operator
(cuda/cuprym_deformation.cu:2554-2855): augmented G = [F0 | -b] and replacement
G = [F0 without two base columns | Z0].

CPU tests pin the oracle's variants against the genus-12 certificate of the
reference's own fixture (tests/fixtures/g12-reference.json):
  * augmented sectors 1 and 5: with b_c = F0_c S_c (S_by_sector, the first-order
    corrections, satisfy blocks[c] @ s_c == b_c in tests/deformation_reference.py:
    195-206) the kernel of [F0_c | -b_c] contains [S_c; I_2], and its rank is
    augmented_shape[1] - augmented_nullity = 84;
  * replacement sector 0: [F0_0 without columns J | Z0] (J = the pivot rows of
    K0^T, tests/deformation_reference.py:211-215) has completion_shape
    [106, 84] and completion_rank 84;
  * both: <G x, y> = <x, G^T y> with the reference's preconditioners.
GPU tests check the library's variants against the oracle, bit for bit.
"""
import numpy as np
import pytest

import oracle_ctypes as oc


def _dense_of(op, p):
    op.set_preconditioners(np.ones(op.rows, np.uint32), np.ones(op.cols, np.uint32))
    eye = np.eye(op.cols, dtype=np.uint32)
    return op.apply_forward(eye.reshape(-1), op.cols).reshape(op.rows, op.cols).astype(np.int64)


def augmented_block(g12, g12_reference, sector):
    """b_c = F0_c S_c mod p (target rows x 2)."""
    p = g12["modulus"]
    F0 = _dense_of(oc.cyclic_op(g12, sector), p)
    S = np.asarray(g12_reference["certificate"]["S_by_sector"][str(sector)], dtype=np.int64)
    return (F0 @ S) % p, S


def replacement_parts(g12_reference):
    cert = g12_reference["certificate"]
    Z0 = np.asarray(cert["Z0"], dtype=np.int64)
    J = list(g12_reference["kernel_pivot_rows_sector0"])
    return Z0, J


@pytest.mark.parametrize("sector", [1, 5])
def test_augmented_kernel_and_rank_match_certificate(g12, g12_reference, sector):
    p = g12["modulus"]
    b, S = augmented_block(g12, g12_reference, sector)
    corr = {c["sector"]: c for c in g12_reference["first_order_corrections"]}[sector]
    assert b.any()  # a sector with a first-order correction
    op = oc.cyclic_op(g12, sector).deform("augmented", b)
    assert [op.rows, op.cols] == corr["augmented_shape"]
    G = _dense_of(op, p)
    # [S_c; I_2] spans the kernel the reference normalises to
    kern = np.vstack([S, np.eye(2, dtype=np.int64)])
    assert not ((G @ kern) % p).any()
    assert oc.rank_mod_p(G.astype(np.uint32), p) == op.cols - corr["augmented_nullity"]


def test_replacement_completion_rank_matches_certificate(g12, g12_reference):
    p = g12["modulus"]
    Z0, J = replacement_parts(g12_reference)
    op = oc.cyclic_op(g12, 0).deform("replacement", Z0, drops=J)
    assert [op.rows, op.cols] == g12_reference["completion_shape"]
    G = _dense_of(op, p)
    assert oc.rank_mod_p(G.astype(np.uint32), p) == g12_reference["completion_rank"]
    # the dense block is the trailing columns, the kept base columns keep their order
    assert (G[:, -2:] == Z0).all()
    base = _dense_of(oc.cyclic_op(g12, 0), p)
    kept = [c for c in range(base.shape[1]) if c not in J]
    assert (G[:, :-2] == base[:, kept]).all()


@pytest.mark.parametrize("variant,sector", [("augmented", 1), ("augmented", 5), ("replacement", 0)])
def test_deformation_transpose_identity_and_square(g12, g12_reference, variant, sector):
    p = g12["modulus"]
    if variant == "augmented":
        dense, _ = augmented_block(g12, g12_reference, sector)
        drops = (0, 0)
    else:
        dense, drops = replacement_parts(g12_reference)
    op = oc.cyclic_op(g12, sector).deform(variant, dense, drops)
    rows, cols, b = op.rows, op.cols, 3
    dr, dc, _ = oc.reference_random_state(1, sector, p, rows, cols, b)
    _, pairs, _ = oc.rowmix_plan(10001, variant, sector, rows, p, 2)
    op.set_preconditioners(dr, dc, 2, pairs)
    rng = np.random.default_rng(sector + 7)
    x = rng.integers(0, p, (cols, b)).astype(np.uint32)
    y = rng.integers(0, p, (rows, b)).astype(np.uint32)
    fx = op.apply_forward(x, b).astype(np.int64).reshape(rows, b)
    fty = op.apply_backward(y, b).astype(np.int64).reshape(cols, b)
    assert ((fx * y.astype(np.int64)).sum() - (x.astype(np.int64) * fty).sum()) % p == 0
    y2 = rng.integers(0, p, (cols, b)).astype(np.uint32)
    sx = op.apply_square(x, b).astype(np.int64).reshape(cols, b)
    sy = op.apply_square(y2, b).astype(np.int64).reshape(cols, b)
    assert ((sx * y2.astype(np.int64)).sum() - (x.astype(np.int64) * sy).sum()) % p == 0


def test_invalid_deformation_requests_are_refused(g12):
    op = oc.cyclic_op(g12, 0)
    dense = np.zeros((op.rows, 2), np.uint32)
    L = oc.lib()
    assert L.orc_cyclic_deformation(op.h, 2, dense.reshape(-1), 2, 1, 1) == -1   # drops not increasing
    assert L.orc_cyclic_deformation(op.h, 2, dense.reshape(-1), 2, 0, op.cols) == -1  # past the base
    assert L.orc_cyclic_deformation(op.h, 3, dense.reshape(-1), 2, 0, 1) == -1   # unknown variant


# ------------------------------------------------------------ C ABI (host) --

def _variant_setup(g12, g12_reference, variant, sector):
    if variant == "augmented":
        dense, _ = augmented_block(g12, g12_reference, sector)
        return dense, (0, 0)
    Z0, J = replacement_parts(g12_reference)
    return Z0, tuple(J)


def test_deformation_dims_and_descriptor_checks(g12, g12_reference):
    import prym_green_ancillary_amd as pg
    base = oc.cyclic_op(g12, 1)
    dense, _ = augmented_block(g12, g12_reference, 1)
    d, keep = pg.Operator.deformation_desc(g12, 1, 2, "augmented", dense)
    assert pg.Operator.dims(d) == (base.rows, base.cols + 2)
    Z0, J = replacement_parts(g12_reference)
    base0 = oc.cyclic_op(g12, 0)
    d, keep = pg.Operator.deformation_desc(g12, 0, 2, "replacement", Z0, (J[1], J[0]))
    assert pg.Operator.dims(d) == (base0.rows, base0.cols)  # kept columns + 2 dense ones
    for drops, msg in [((3, 3), "distinct"), ((0, base0.cols), "outside the base")]:
        d, keep = pg.Operator.deformation_desc(g12, 0, 2, "replacement", Z0, drops)
        with pytest.raises(pg.PgwError, match=msg):
            pg.Operator.dims(d)
    d, keep = pg.Operator.deformation_desc(g12, 0, 2, "augmented", np.zeros((base0.rows, 3), np.uint32))
    with pytest.raises(pg.PgwError, match="two columns"):
        pg.Operator.dims(d)


# --------------------------------------------------------------------- GPU --

@pytest.fixture(scope="module")
def gctx():
    import torch
    import prym_green_ancillary_amd as pg
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    c = pg.Context(0)
    yield c
    c.close()


def _pair(g12, g12_reference, gctx, variant, sector, b, rounds=2):
    import prym_green_ancillary_amd as pg
    p = g12["modulus"]
    dense, drops = _variant_setup(g12, g12_reference, variant, sector)
    o = oc.cyclic_op(g12, sector).deform(variant, dense, drops)
    dr, dc, v = pg.reference_random_state(1, sector, p, o.rows, o.cols, b)
    pairs = pg.rowmix_plan(pg.rowmix_realized_seed(10001, variant, sector, o.rows, p), o.rows, rounds, p)
    o.set_preconditioners(dr, dc, rounds, pairs.reshape(-1))
    op = pg.Operator.deformation(gctx, g12, sector, b, variant, dense, drops, dr, dc, rounds, pairs)
    assert (op.rows, op.cols) == (o.rows, o.cols)
    return o, op, dr, dc, v


@pytest.mark.gpu
@pytest.mark.parametrize("variant,sector,b", [("augmented", 1, 4), ("augmented", 5, 3),
                                              ("replacement", 0, 4), ("replacement", 0, 1)])
def test_gpu_deformation_apply_matches_oracle(g12, g12_reference, gctx, variant, sector, b):
    import torch
    print(f"deformation apply {variant} sector {sector} b={b}", flush=True)
    o, op, dr, dc, v = _pair(g12, g12_reference, gctx, variant, sector, b)
    x = torch.from_numpy(v.view(np.int32).copy()).cuda()
    y = torch.empty_like(x)
    op.apply_square(x.data_ptr(), y.data_ptr(), torch.cuda.current_stream().cuda_stream)
    t = torch.zeros(o.rows * b, dtype=torch.int32, device="cuda")
    op.apply_forward(x.data_ptr(), t.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert (y.cpu().numpy().view(np.uint32) == o.apply_square(v, b)).all()
    assert (t.cpu().numpy().view(np.uint32) == o.apply_forward(v, b)).all()


@pytest.mark.gpu
@pytest.mark.parametrize("variant,sector", [("augmented", 5), ("replacement", 0)])
def test_gpu_deformation_sequence_and_kernel_vectors_match_oracle(g12, g12_reference, gctx, variant, sector):
    b = 4
    print(f"deformation sequence {variant} sector {sector}", flush=True)
    o, op, dr, dc, v = _pair(g12, g12_reference, gctx, variant, sector, b)
    max_len = 2 * o.cols // b + 50  # the reference's default length (:2858)
    seq, curv = op.sequence(v, max_len)
    oseq, ocurv = o.sequence(v, b, max_len)
    assert (seq == oseq).all() and (curv == ocurv).all()
    gens = np.random.default_rng(sector).integers(0, g12["modulus"], size=(7, b * b)).astype(np.uint32)
    d, prod = op.kernel_vectors(v, gens)
    od, oprod = o.kernel_vectors(v, b, gens)
    assert (d == od).all() and (prod == oprod).all()


@pytest.mark.gpu
@pytest.mark.parametrize("sector", [1, 5])
def test_gpu_augmented_kernel_certificate(g12, g12_reference, gctx, sector):
    # G Dc (Dc^-1 [S_c; I]) = [F0 | -b_c] [S_c; I] = F0 S_c - b_c = 0 on the GPU
    import torch
    import prym_green_ancillary_amd as pg
    p, b = g12["modulus"], 2
    dense, S = augmented_block(g12, g12_reference, sector)
    o = oc.cyclic_op(g12, sector).deform("augmented", dense)
    dr, dc, _ = pg.reference_random_state(1, sector, p, o.rows, o.cols, b)
    op = pg.Operator.deformation(gctx, g12, sector, b, "augmented", dense, (0, 0), dr, dc, 0, None)
    kern = np.vstack([S, np.eye(2, dtype=np.int64)])
    inv = np.array([pow(int(c), p - 2, p) for c in dc], dtype=np.int64)
    x = torch.from_numpy(((kern * inv[:, None]) % p).astype(np.uint32).view(np.int32).copy()).cuda()
    t = torch.ones(o.rows * b, dtype=torch.int32, device="cuda")
    op.apply_forward(x.data_ptr(), t.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert not t.cpu().numpy().any()


# --------------------------------------------------------------------- CLI --

def _write_inputs(tmp_path, p, dense, drops=None):
    dense = np.asarray(dense, dtype=np.int64)
    df = tmp_path / "dense.json"
    df.write_text(json_dumps({"rows": dense.shape[0], "columns": dense.shape[1], "prime": p,
                              "values": dense.tolist()}))
    if drops is None:
        return df, None
    dc = tmp_path / "drop.json"
    dc.write_text(json_dumps({"indices": list(drops)}))
    return df, dc


def json_dumps(obj):
    import json
    return json.dumps(obj)


def _cli(args):
    import subprocess
    gen = oc.ROOT / "prym_green_ancillary_amd" / "bin" / "cuprym_cyclic"
    return subprocess.run([str(gen)] + [str(a) for a in args], capture_output=True, text=True, timeout=600)


G12 = oc.ROOT / "tests" / "golden" / "g12.json"


def test_generator_cli_refuses_bad_deformation_inputs(tmp_path, g12, g12_reference):
    p = g12["modulus"]
    Z0, J = replacement_parts(g12_reference)
    df, dcf = _write_inputs(tmp_path, p, Z0, J)
    cases = [
        (["--operator", "replacement", "--dense-file", df], "requires --drop-columns"),
        (["--operator", "augmented", "--dense-file", df, "--drop-columns", dcf], "only meaningful for replacement"),
        (["--operator", "augmented", "--dense-file", df, "--sector", 1], "dimensions do not match"),
    ]
    bad_drop = tmp_path / "bad.json"
    bad_drop.write_text('{"indices": [3, 3]}')
    cases.append((["--operator", "replacement", "--dense-file", df, "--drop-columns", bad_drop], "distinct"))
    wrong_p = tmp_path / "wrongp.json"
    wrong_p.write_text(json_dumps({"rows": Z0.shape[0], "columns": 2, "prime": 661, "values": Z0.tolist()}))
    cases.append((["--operator", "augmented", "--dense-file", wrong_p], "prime does not match"))
    for extra, msg in cases:
        args = [G12, "--no-save", "--out", tmp_path] + extra
        if "--sector" not in extra:
            args += ["--sector", 0]
        r = _cli(args)
        assert r.returncode == 1 and msg in (r.stderr + r.stdout), (extra, r.stderr)


@pytest.mark.gpu
@pytest.mark.parametrize("variant,sector", [("augmented", 1), ("replacement", 0)])
def test_generator_cli_deformation_writes_oracle_wdm(tmp_path, g12, g12_reference, variant, sector):
    import test_cli as tc
    p, b = g12["modulus"], 4
    dense, drops = _variant_setup(g12, g12_reference, variant, sector)
    df, dcf = _write_inputs(tmp_path, p, dense, drops if variant == "replacement" else None)
    args = [G12, "--sector", sector, "--out", tmp_path, "-v", b, "--operator", variant, "--dense-file", df]
    if dcf is not None:
        args += ["--drop-columns", dcf]
    r = _cli(args)
    assert r.returncode == 0, r.stderr
    name = f"{variant}-sector{sector:02d}-v{b}-rowmix-r2-s10001"
    text = tc.decompress(tmp_path / f"{name}.wdm.zst")
    o = oc.cyclic_op(g12, sector).deform(variant, dense, drops)
    dr, dc, v = oc.reference_random_state(1, sector, p, o.rows, o.cols, b)
    _, pairs, _ = oc.rowmix_plan(10001, variant, sector, o.rows, p, 2)
    o.set_preconditioners(dr, dc, 2, pairs)
    seq, curv = o.sequence(v, b, 2 * o.cols // b + 50)
    ref = tmp_path / "oracle.wdm"
    oc.lib().orc_write_wdm(str(ref).encode(), o.rows, o.cols, p, dr, dc, v, curv, b,
                           np.ascontiguousarray(seq.reshape(-1)), seq.shape[0])
    assert text == ref.read_text()


@pytest.mark.gpu
@pytest.mark.parametrize("variant,sector,rank", [
    ("augmented", 1, 84), ("replacement", 0, 84),
    # over F_109 the reference's fixed preconditioners lose one Gram rank in
    # augmented sector 5 (tests/check_cuda.py:71-73); the same recipe here
    # reproduces that loss, the exact rank (86 - 2) being pinned above
    ("augmented", 5, 83)])
def test_gpu_deformation_sequence_rank(g12, g12_reference, gctx, variant, sector, rank):
    import prym_green_ancillary_amd as pg
    p, b = g12["modulus"], 4
    print(f"deformation rank {variant} sector {sector}", flush=True)
    o, op, dr, dc, v = _pair(g12, g12_reference, gctx, variant, sector, b)
    seq, _ = op.sequence(v, 2 * o.cols // b + 50)
    r = pg.RankComputation(p, b, o.rows, o.cols)
    st = r.step(seq)
    assert st.status == pg.RANK_SUCCESS and st.rank_lower_bound == rank
    _, rank_o, spread_o, gen_o = oc.bcw_rank(p, b, seq)
    assert (rank_o, spread_o) == (st.rank_lower_bound, st.delta_spread)
    assert np.array_equal(r.generator(), gen_o)
    # the generator's kernel vectors are annihilated by G (the certificate)
    # exactly when no Gram rank was lost; a lost rank leaves a vector in the
    # kernel of Dc G^T P^T Dr P G Dc that G does not annihilate
    d, prod = op.kernel_vectors(v, r.generator().reshape(-1, b * b))
    exact = o.cols - 2 if variant == "augmented" else g12_reference["completion_rank"]
    assert bool(prod.any()) == (rank != exact)


@pytest.mark.gpu
@pytest.mark.parametrize("variant,sector", [("augmented", 5), ("replacement", 0)])
def test_generator_cli_deformation_validate_rowmix(tmp_path, g12, g12_reference, variant, sector):
    # the device apply of a deformation variant equals the exact host apply
    print(f"validate-rowmix {variant} sector {sector}", flush=True)
    p = g12["modulus"]
    dense, drops = _variant_setup(g12, g12_reference, variant, sector)
    df, dcf = _write_inputs(tmp_path, p, dense, drops if variant == "replacement" else None)
    args = [G12, "--sector", sector, "--out", tmp_path, "-v", 4, "--operator", variant, "--dense-file", df,
            "--validate-rowmix", "--no-save", "--length", 9]
    if dcf is not None:
        args += ["--drop-columns", dcf]
    r = _cli(args)
    assert r.returncode == 0, r.stderr[-2000:]
    for idx in range(3):
        assert f"Validation matched: {variant} square block {idx} (" in r.stdout, r.stdout[-2000:]

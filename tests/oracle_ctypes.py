"""ctypes binding of the CPU oracle (oracle/pgw_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product path never
imports this module.
"""
from __future__ import annotations

import ctypes
import json
import os
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
ORACLE_SO = ORACLE_DIR / "_build" / "libpgw_oracle.so"

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_c = ctypes
_lib = None


def build_oracle() -> Path:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return ORACLE_SO


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not ORACLE_SO.exists():
        build_oracle()
    L = ctypes.CDLL(str(ORACLE_SO))
    sz = _c.c_size_t
    L.orc_binom.restype = _c.c_uint64
    L.orc_binom.argtypes = [_c.c_uint32, _c.c_uint32]
    L.orc_unrank_subset.restype = _c.c_int
    L.orc_unrank_subset.argtypes = [_c.c_uint64, _c.c_uint32, _c.c_uint32, _u32p]
    L.orc_splitmix64_next.restype = _c.c_uint64
    L.orc_splitmix64_next.argtypes = [_c.POINTER(_c.c_uint64)]
    L.orc_seeded_sector_value.restype = _c.c_uint64
    L.orc_seeded_sector_value.argtypes = [_c.c_uint64, _c.c_uint64]
    L.orc_fnv1a64_u32_vector.restype = _c.c_uint64
    L.orc_fnv1a64_u32_vector.argtypes = [_u32p, sz]
    L.orc_fnv1a64_string.restype = _c.c_uint64
    L.orc_fnv1a64_string.argtypes = [_c.c_char_p]
    L.orc_mt19937_64_outputs.restype = None
    L.orc_mt19937_64_outputs.argtypes = [_c.c_uint64, sz, _u64p]
    L.orc_uniform_u64.restype = None
    L.orc_uniform_u64.argtypes = [_c.c_uint64, _c.c_uint64, _c.c_uint64, sz, _u64p]
    L.orc_reference_random_state.restype = None
    L.orc_reference_random_state.argtypes = [_c.c_uint64, _c.c_uint64, _c.c_uint32, sz, sz, sz,
                                             _u32p, _u32p, _u32p]
    L.orc_rowmix_realized_seed.restype = _c.c_uint64
    L.orc_rowmix_realized_seed.argtypes = [_c.c_uint64, _c.c_char_p, _c.c_uint64, _c.c_uint64,
                                           _c.c_uint32]
    L.orc_rowmix_plan.restype = None
    L.orc_rowmix_plan.argtypes = [_c.c_uint64, sz, sz, _c.c_uint32, _u32p]
    L.orc_rowmix_plan_hash.restype = _c.c_uint64
    L.orc_rowmix_plan_hash.argtypes = [sz, sz, _c.c_uint64, _c.c_uint64, _u32p]
    L.orc_synthetic_table.restype = None
    L.orc_synthetic_table.argtypes = [_c.c_uint64, _c.c_uint32, sz, _u32p]
    L.orc_synthetic_create.restype = _c.c_void_p
    L.orc_synthetic_create.argtypes = [_c.c_uint32] * 5 + [_u32p, _u32p, _u32p, _u32p, sz, _c.c_void_p]
    L.orc_cyclic_create.restype = _c.c_void_p
    L.orc_cyclic_create.argtypes = [_c.c_uint32] * 4 + [_u32p, _u16p, _u16p, _u16p, _c.c_uint32,
                                                        _c.c_uint32, _c.c_uint32, _u16p, _u32p, _u32p]
    L.orc_cyclic_deformation.restype = _c.c_int
    L.orc_cyclic_deformation.argtypes = [_c.c_void_p, _c.c_int, _u32p, sz, sz, sz]
    L.orc_set_preconditioners.restype = None
    L.orc_set_preconditioners.argtypes = [_c.c_void_p, _u32p, _u32p, sz, _c.c_void_p]
    L.orc_op_dims.restype = None
    L.orc_op_dims.argtypes = [_c.c_void_p, _c.POINTER(sz), _c.POINTER(sz)]
    L.orc_op_destroy.restype = None
    L.orc_op_destroy.argtypes = [_c.c_void_p]
    for name in ("orc_apply_forward", "orc_apply_backward", "orc_apply_square"):
        f = getattr(L, name)
        f.restype = None
        f.argtypes = [_c.c_void_p, _u32p, _u32p, sz]
    L.orc_sweep_partial.restype = None
    L.orc_sweep_partial.argtypes = [_c.c_void_p, _c.c_int, _u32p, _u32p, sz, sz, sz]
    L.orc_elementwise_sample.restype = None
    L.orc_elementwise_sample.argtypes = [_c.c_void_p, _u32p, sz, sz]
    L.orc_set_threads.restype = None
    L.orc_set_threads.argtypes = [_c.c_int]
    L.orc_sweep_rows.restype = sz
    L.orc_sweep_rows.argtypes = [_c.c_void_p, _c.c_int]
    L.orc_gram_upper.restype = None
    L.orc_gram_upper.argtypes = [_u32p, _u32p, sz, sz, _c.c_uint32, _u32p]
    L.orc_sequence.restype = sz
    L.orc_sequence.argtypes = [_c.c_void_p, _u32p, sz, sz, _u32p, _u32p]
    L.orc_accumulate_generator.restype = None
    L.orc_accumulate_generator.argtypes = [_u32p, _u32p, _u32p, sz, sz, _c.c_uint32]
    L.orc_kernel_vectors.restype = None
    L.orc_kernel_vectors.argtypes = [_c.c_void_p, _u32p, sz, _u32p, sz, _u32p, _u32p]
    L.orc_rank_mod_p.restype = sz
    L.orc_rank_mod_p.argtypes = [_u32p, sz, sz, _c.c_uint32, _c.c_void_p]
    L.orc_write_wdm.restype = _c.c_int
    L.orc_write_wdm.argtypes = [_c.c_char_p, sz, sz, _c.c_uint32, _u32p, _u32p, _u32p, _u32p, sz,
                                _u32p, sz]
    L.orc_bcw_rank.restype = _c.c_int
    L.orc_bcw_rank.argtypes = [_c.c_uint32, sz, _u32p, sz, _c.c_void_p, _c.POINTER(_c.c_uint64),
                               _c.POINTER(_c.c_uint64), _c.c_void_p, sz, _c.POINTER(sz)]
    _lib = L
    return L


def u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


class OracleOp:
    """Owns an oracle operator plus the arrays it points into."""

    def __init__(self, handle, keep, p):
        self.h = handle
        self._keep = list(keep)
        self.p = p
        rows, cols = ctypes.c_size_t(), ctypes.c_size_t()
        lib().orc_op_dims(self.h, ctypes.byref(rows), ctypes.byref(cols))
        self.rows, self.cols = rows.value, cols.value

    def deform(self, variant: str, dense, drops=(0, 0)):
        """Turn an eliminated cyclic operator into its augmented [F0 | -dense]
        or replacement [F0 without the two drop columns | dense] variant."""
        dense = u32(dense)
        assert dense.ndim == 2 and dense.shape[0] == self.rows
        code = {"augmented": 1, "replacement": 2}[variant]
        self._keep.append(dense)
        rc = lib().orc_cyclic_deformation(self.h, code, dense.reshape(-1), dense.shape[1],
                                          drops[0], drops[1])
        assert rc == 0, "invalid deformation request"
        rows, cols = ctypes.c_size_t(), ctypes.c_size_t()
        lib().orc_op_dims(self.h, ctypes.byref(rows), ctypes.byref(cols))
        self.rows, self.cols = rows.value, cols.value
        return self

    def set_preconditioners(self, dr, dc, rounds=0, pairs=None):
        dr, dc = u32(dr), u32(dc)
        pairs_arr = u32(pairs) if pairs is not None else None
        self._keep += [dr, dc, pairs_arr]
        lib().orc_set_preconditioners(self.h, dr, dc, rounds,
                                      pairs_arr.ctypes.data if pairs_arr is not None else None)

    def apply_square(self, x, b):
        x = u32(x).reshape(-1)
        out = np.zeros(self.cols * b, dtype=np.uint32)
        lib().orc_apply_square(self.h, x, out, b)
        return out

    def apply_forward(self, x, b):
        x = u32(x).reshape(-1)
        out = np.zeros(self.rows * b, dtype=np.uint32)
        lib().orc_apply_forward(self.h, x, out, b)
        return out

    def apply_backward(self, y, b):
        y = u32(y).reshape(-1)
        out = np.zeros(self.cols * b, dtype=np.uint32)
        lib().orc_apply_backward(self.h, y, out, b)
        return out

    def sequence(self, v, b, max_len):
        v = u32(v).reshape(-1)
        ut = b * (b + 1) // 2
        L = 1 + 4 * ((max_len - 1 + 3) // 4)
        seq = np.zeros(L * ut, dtype=np.uint32)
        curv = np.zeros(self.cols * b, dtype=np.uint32)
        got = lib().orc_sequence(self.h, v, b, max_len, seq, curv)
        assert got == L
        return seq.reshape(L, ut), curv

    def kernel_vectors(self, v, b, gens):
        v = u32(v).reshape(-1)
        gens = u32(gens).reshape(-1)
        ngen = gens.size // (b * b)
        d = np.zeros(self.cols * b, dtype=np.uint32)
        prod = np.zeros(self.rows * b, dtype=np.uint32)
        lib().orc_kernel_vectors(self.h, v, b, gens, ngen, d, prod)
        return d, prod

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_op_destroy(self.h)
            self.h = None


def synthetic_op(n, k, M, Mp, p, T, E, dr, dc, rounds=0, pairs=None):
    T, E, dr, dc = u32(T), u32(E), u32(dr), u32(dc)
    pairs_arr = u32(pairs) if pairs is not None else None
    h = lib().orc_synthetic_create(n, k, M, Mp, p, T, E, dr, dc, rounds,
                                   pairs_arr.ctypes.data if pairs_arr is not None else None)
    return OracleOp(h, [T, E, dr, dc, pairs_arr], p)


def cyclic_op(inst: dict, sector: int):
    """Eliminated cyclic operator from a parsed prym-cyclic instance JSON."""
    L = lib()
    g, p, r = inst["genus"], inst["modulus"], inst["cyclic"]["order"]
    n = g - 3
    mu = u32(np.asarray(inst["artinian"]["mu"]["values"]).reshape(-1))
    w = inst["weights"]
    wv = np.ascontiguousarray(np.asarray(w["V"], dtype=np.uint16))
    wa0 = np.ascontiguousarray(np.asarray(w["A0"], dtype=np.uint16))
    wa1 = np.ascontiguousarray(np.asarray(w["A1"], dtype=np.uint16))
    e = inst["third_section_elimination"]
    kw = np.ascontiguousarray(np.asarray(e["kernel_weights"], dtype=np.uint16))
    R = u32(np.asarray(e["right_inverse_A0_by_A1"]).reshape(-1))
    B = u32(np.asarray(e["kernel_basis_A0_by_3"]).reshape(-1))
    assert mu.size == n * n * g
    h = L.orc_cyclic_create(g, r, p, sector, mu, wv, wa0, wa1, e["w_V_index"], e["w_weight"],
                            kw.size, kw, R, B)
    return OracleOp(h, [mu, wv, wa0, wa1, kw, R, B], p)


def reference_random_state(seed, sector, p, rows, cols, b):
    dr = np.zeros(rows, dtype=np.uint32)
    dc = np.zeros(cols, dtype=np.uint32)
    v = np.zeros(cols * b, dtype=np.uint32)
    lib().orc_reference_random_state(seed, sector, p, rows, cols, b, dr, dc, v)
    return dr, dc, v


def rowmix_plan(seed, op_name, sector, rows, p, rounds):
    L = lib()
    realized = L.orc_rowmix_realized_seed(seed, op_name.encode(), sector, rows, p)
    pairs = np.zeros(rounds * (rows // 2) * 4 + 1, dtype=np.uint32)
    L.orc_rowmix_plan(realized, rows, rounds, p, pairs)
    pairs = pairs[:-1].copy()
    h = L.orc_rowmix_plan_hash(rows, rounds, seed, realized, pairs if pairs.size else np.zeros(1, np.uint32))
    return realized, pairs, h


def synthetic_table(seed, p, count):
    out = np.zeros(count, dtype=np.uint32)
    lib().orc_synthetic_table(seed, p, count, out)
    return out


def gram_upper(x, y, rows, b, p):
    out = np.zeros(b * (b + 1) // 2, dtype=np.uint32)
    lib().orc_gram_upper(u32(x).reshape(-1), u32(y).reshape(-1), rows, b, p, out)
    return out


def rank_mod_p(m, p):
    m = u32(m)
    rows, cols = m.shape
    return lib().orc_rank_mod_p(m.reshape(-1), rows, cols, p, None)


def bcw_rank(p, b, seq):
    """Iterative BCW restatement: (delta[2b], rank, spread, generator[terms, b, b])."""
    seq = np.ascontiguousarray(u32(seq).reshape(-1))
    L = seq.size // (b * (b + 1) // 2)
    delta = np.zeros(2 * b, dtype=np.uint64)
    rank, spread, terms = _c.c_uint64(), _c.c_uint64(), _c.c_size_t()
    rc = lib().orc_bcw_rank(p, b, seq, L, delta.ctypes.data, _c.byref(rank), _c.byref(spread),
                            None, 0, _c.byref(terms))
    assert rc == 0, rc
    gen = np.zeros(terms.value * b * b, dtype=np.uint32)
    rc = lib().orc_bcw_rank(p, b, seq, L, delta.ctypes.data, _c.byref(rank), _c.byref(spread),
                            gen.ctypes.data, gen.size, _c.byref(terms))
    assert rc == 0, rc
    return delta, rank.value, spread.value, gen.reshape(terms.value, b, b)


def load_json(path):
    with open(path) as f:
        return json.load(f)

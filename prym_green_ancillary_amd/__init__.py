"""MI355X block Wiedemann sequence generator -- Python host interface.

A thin ctypes layer over libpgw.so (include/pgw.h).  It mirrors the host side
of the reference's `cuprym_cyclic` driver (cuda/cuprym_cyclic.cu): the
operator of one sector, its preconditioners and row-mixing plan, the block
apply S = Dc F^T P^T Dr P F Dc, the packed Gram projection and the whole
sequence loop.  All arithmetic runs in the HIP library; this module never
computes a residue itself, and it raises if the native library is missing.
"""
from __future__ import annotations

import ctypes
import json
from dataclasses import dataclass
from pathlib import Path

import numpy as np

__all__ = [
    "NativeLibraryError", "PgwError", "lib", "library_path", "Context", "Operator",
    "reference_random_state", "rowmix_realized_seed", "rowmix_plan", "rowmix_plan_hash",
    "deterministic_block", "cyclic_apply_cpu",
    "fnv1a64_u32", "synthetic_table", "SyntheticSpec", "synthetic_spec_from_json",
    "load_cyclic_instance", "METHOD_AUTO", "METHOD_PLAIN", "METHOD_MFMA", "sequence_length",
    "RankComputation", "RANK_RUNNING", "RANK_SUCCESS", "RANK_EXCEEDS_DIMENSION",
    "RANK_SEQUENCE_LIMIT", "NEED_MORE_SEQUENCE",
]

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libpgw.so"

# pgw_rank_status (RankComputationStatus) and the need-more-sequence error
RANK_RUNNING, RANK_SUCCESS, RANK_EXCEEDS_DIMENSION, RANK_SEQUENCE_LIMIT = 0, 1, 2, 3
NEED_MORE_SEQUENCE = 6
DEFAULT_STOPPING_THRESHOLD = 10
DEFAULT_MAX_NLEN_BUFFER = 40

KIND_SYNTHETIC = 1
KIND_CYCLIC = 2
KIND_CYCLIC_AUGMENTED = 3
KIND_CYCLIC_REPLACEMENT = 4
METHOD_AUTO, METHOD_PLAIN, METHOD_MFMA = 0, 1, 2


class NativeLibraryError(RuntimeError):
    """libpgw.so is missing or does not load: there is no fallback path."""


class PgwError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"pgw error {status}: {message}")
        self.status = status


class RowMixPair(ctypes.Structure):
    _fields_ = [("first", ctypes.c_uint32), ("second", ctypes.c_uint32),
                ("a", ctypes.c_uint32), ("b", ctypes.c_uint32)]


class OperatorDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_uint32), ("prime", ctypes.c_uint32), ("block", ctypes.c_uint32),
        ("method", ctypes.c_uint32),
        ("n", ctypes.c_uint32), ("k", ctypes.c_uint32), ("m_src", ctypes.c_uint32),
        ("m_tgt", ctypes.c_uint32),
        ("table_d", ctypes.c_void_p), ("table_e", ctypes.c_void_p),
        ("genus", ctypes.c_uint32), ("order", ctypes.c_uint32), ("sector", ctypes.c_uint32),
        ("mu", ctypes.c_void_p), ("weight_v", ctypes.c_void_p), ("weight_a0", ctypes.c_void_p),
        ("weight_a1", ctypes.c_void_p),
        ("w_index", ctypes.c_uint32), ("w_weight", ctypes.c_uint32), ("kernel_dim", ctypes.c_uint32),
        ("kernel_weights", ctypes.c_void_p), ("right_inverse", ctypes.c_void_p),
        ("kernel_basis", ctypes.c_void_p),
        ("dense_cols", ctypes.c_uint32), ("dense_block", ctypes.c_void_p),
        ("drop_cols", ctypes.c_uint32 * 2),
    ]


class RankInfo(ctypes.Structure):
    _fields_ = [("rank", ctypes.c_uint64), ("delta_spread", ctypes.c_uint64),
                ("processed", ctypes.c_uint64), ("max_terms", ctypes.c_uint64),
                ("status", ctypes.c_int)]


class OpInfo(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_uint64), ("cols", ctypes.c_uint64), ("mid_rows", ctypes.c_uint64),
                ("device_bytes", ctypes.c_uint64), ("method", ctypes.c_uint32),
                ("block", ctypes.c_uint32)]


PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p)

_lib = None

# every symbol include/pgw.h declares (checked by the CPU test suite)
EXPORTED = [
    "pgw_last_error", "pgw_ctx_create", "pgw_ctx_destroy", "pgw_ctx_synchronize",
    "pgw_ctx_device_info", "pgw_seeded_sector_value", "pgw_reference_random_state",
    "pgw_rowmix_realized_seed", "pgw_rowmix_plan", "pgw_rowmix_plan_hash", "pgw_fnv1a64_u32",
    "pgw_synthetic_table", "pgw_operator_dims", "pgw_operator_create", "pgw_operator_destroy",
    "pgw_operator_info", "pgw_apply_square", "pgw_apply_forward", "pgw_gram_upper",
    "pgw_sequence_run", "pgw_kernel_vectors", "pgw_bench_apply", "pgw_operator_autotune",
    "pgw_batch_create", "pgw_batch_add", "pgw_batch_step", "pgw_batch_poll", "pgw_batch_result",
    "pgw_batch_cancel", "pgw_batch_graph_builds", "pgw_batch_groups", "pgw_batch_destroy", "pgw_server_create",
    "pgw_server_submit", "pgw_server_wait", "pgw_server_result", "pgw_server_stats",
    "pgw_server_destroy", "pgw_job_footprint", "pgw_rank_create", "pgw_rank_step",
    "pgw_rank_get_info", "pgw_rank_delta", "pgw_rank_generator", "pgw_rank_destroy",
    "pgw_server_enable_rank", "pgw_server_rank", "pgw_wdm_write", "pgw_bench_stages",
    "pgw_gram_upper_noalloc", "pgw_cyclic_apply_cpu", "pgw_deterministic_block",
    "pgw_wdm_stream_open", "pgw_wdm_stream_terms", "pgw_wdm_stream_finish", "pgw_wdm_stream_destroy",
    "pgw_operator_set_gram", "pgw_batch_terms", "pgw_operator_prepare_row_mixing",
]


def library_path() -> Path:
    """The in-tree libpgw.so; no other build is ever loaded."""
    return LIB_PATH


def lib():
    """Load libpgw.so (built in-tree by __graft_entry__.build()).

    PyTorch is imported first when it is installed: its wheel carries its own
    HIP and HSA runtimes, which libpgw then shares (same soname).  Loaded
    the other way round, libpgw would pull in /opt/rocm's runtimes and a
    later `import torch` a second pair beside them, two runtimes driving one
    GPU in one process."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        import torch  # noqa: F401  (runtime load order, see above)
    except ImportError:
        pass
    path = library_path()
    if not path.exists():
        raise NativeLibraryError(f"{path} is missing: build it with __graft_entry__.build() "
                                 "(make -C prym_green_ancillary_amd/csrc)")
    try:
        L = ctypes.CDLL(str(path))
    except OSError as e:  # pragma: no cover - loader failure
        raise NativeLibraryError(f"cannot load {path}: {e}") from e
    u32, u64, vp, sz = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_size_t
    I = ctypes.c_int
    sig = {
        "pgw_last_error": (ctypes.c_char_p, []),
        "pgw_ctx_create": (I, [I, ctypes.POINTER(vp)]),
        "pgw_ctx_destroy": (I, [vp]),
        "pgw_ctx_synchronize": (I, [vp]),
        "pgw_ctx_device_info": (I, [vp, ctypes.c_char_p, sz, ctypes.POINTER(u64)]),
        "pgw_seeded_sector_value": (u64, [u64, u64]),
        "pgw_reference_random_state": (I, [u64, u64, u32, u64, u64, u32, vp, vp, vp]),
        "pgw_rowmix_realized_seed": (u64, [u64, ctypes.c_char_p, u64, u64, u32]),
        "pgw_rowmix_plan": (I, [u64, u64, u32, u32, vp]),
        "pgw_rowmix_plan_hash": (u64, [u64, u32, u64, u64, vp]),
        "pgw_fnv1a64_u32": (u64, [vp, u64]),
        "pgw_synthetic_table": (I, [u64, u32, u64, vp]),
        "pgw_operator_dims": (I, [ctypes.POINTER(OperatorDesc), ctypes.POINTER(u64), ctypes.POINTER(u64)]),
        "pgw_operator_create": (I, [vp, ctypes.POINTER(OperatorDesc), vp, vp, u32, vp, ctypes.POINTER(vp)]),
        "pgw_operator_destroy": (I, [vp]),
        "pgw_operator_info": (I, [vp, ctypes.POINTER(OpInfo)]),
        "pgw_apply_square": (I, [vp, vp, vp, vp]),
        "pgw_apply_forward": (I, [vp, vp, vp, vp]),
        "pgw_gram_upper": (I, [vp, vp, vp, vp, vp]),
        "pgw_gram_upper_noalloc": (I, [vp, u32, u32, u64, vp, vp, vp, vp]),
        "pgw_cyclic_apply_cpu": (I, [ctypes.POINTER(OperatorDesc), vp, vp, u32, vp, vp, u32, vp]),
        "pgw_deterministic_block": (I, [u64, u64, u32, u64, vp]),
        "pgw_sequence_run": (I, [vp, vp, u64, vp, vp, ctypes.POINTER(u64), PROGRESS_FN, vp]),
        "pgw_kernel_vectors": (I, [vp, vp, vp, u64, vp, vp]),
        "pgw_bench_apply": (I, [vp, vp, vp, u32, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u32)]),
        "pgw_operator_autotune": (I, [vp, u32, vp, vp]),
        "pgw_bench_stages": (I, [vp, vp, vp, u32, vp, vp, vp]),
        "pgw_batch_create": (I, [vp, u32, ctypes.POINTER(vp)]),
        "pgw_batch_groups": (I, [vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]),
        "pgw_batch_add": (I, [vp, vp, vp, u64, ctypes.POINTER(u64)]),
        "pgw_batch_step": (I, [vp, u32, ctypes.POINTER(u64)]),
        "pgw_batch_poll": (I, [vp, u64, ctypes.POINTER(I)]),
        "pgw_batch_result": (I, [vp, u64, vp, vp, ctypes.POINTER(u64)]),
        "pgw_batch_cancel": (I, [vp, u64]),
        "pgw_batch_graph_builds": (I, [vp, ctypes.POINTER(u64)]),
        "pgw_batch_destroy": (I, [vp]),
        "pgw_server_create": (I, [vp, u64, u32, u32, ctypes.POINTER(vp)]),
        "pgw_server_submit": (I, [vp, ctypes.POINTER(OperatorDesc), vp, vp, u32, vp, vp, u64,
                                  ctypes.POINTER(u64)]),
        "pgw_server_wait": (I, [vp, u64, ctypes.c_double, ctypes.POINTER(I)]),
        "pgw_server_result": (I, [vp, u64, vp, vp, ctypes.POINTER(u64), vp]),
        "pgw_server_stats": (I, [vp, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64),
                                 ctypes.POINTER(u64)]),
        "pgw_server_destroy": (I, [vp]),
        "pgw_job_footprint": (I, [ctypes.POINTER(OperatorDesc), u64, u32, ctypes.POINTER(u64)]),
        "pgw_rank_create": (I, [u32, u32, u64, u64, u32, u32, u32, ctypes.POINTER(vp)]),
        "pgw_rank_step": (I, [vp, vp, u64, ctypes.POINTER(RankInfo)]),
        "pgw_rank_get_info": (I, [vp, ctypes.POINTER(RankInfo)]),
        "pgw_rank_delta": (I, [vp, vp]),
        "pgw_rank_generator": (I, [vp, vp, u64, ctypes.POINTER(u64)]),
        "pgw_rank_destroy": (I, [vp]),
        "pgw_server_enable_rank": (I, [vp, u32, u32, u32, I]),
        "pgw_server_rank": (I, [vp, u64, ctypes.POINTER(RankInfo), vp, u64, ctypes.POINTER(u64),
                                ctypes.POINTER(ctypes.c_double)]),
        "pgw_wdm_write": (I, [ctypes.c_char_p, u64, u64, u32, vp, vp, vp, vp, u32, vp, u64, u32]),
        "pgw_wdm_stream_open": (I, [ctypes.c_char_p, u64, u64, u32, vp, vp, vp, u32, u64, u32, u64,
                                    ctypes.POINTER(vp)]),
        "pgw_wdm_stream_terms": (I, [vp, vp, u64]),
        "pgw_wdm_stream_finish": (I, [vp, vp]),
        "pgw_wdm_stream_destroy": (I, [vp]),
        "pgw_operator_set_gram": (I, [vp, u32]),
        "pgw_operator_prepare_row_mixing": (I, [vp]),
        "pgw_batch_terms": (I, [vp, u64, u64, vp, ctypes.POINTER(u64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise PgwError(rc, lib().pgw_last_error().decode())


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------------------
# host recipe helpers (cuda/cuprym_cyclic.cu:1143-1287)
# ---------------------------------------------------------------------------

def reference_random_state(seed: int, sector: int, p: int, rows: int, cols: int, b: int):
    """Dr, Dc, V exactly as run_eliminated_sector draws them (cuprym_cyclic.cu:2196-2199)."""
    dr = np.empty(rows, np.uint32)
    dc = np.empty(cols, np.uint32)
    v = np.empty(cols * b, np.uint32)
    _check(lib().pgw_reference_random_state(seed, sector, p, rows, cols, b, _ptr(dr), _ptr(dc), _ptr(v)))
    return dr, dc, v


def deterministic_block(rows: int, cols: int, p: int, salt: int) -> np.ndarray:
    """deterministic_dense_block (cuprym_cyclic.cu:1410-1420), rows x cols row-major."""
    out = np.empty(max(1, rows * cols), np.uint32)
    _check(lib().pgw_deterministic_block(rows, cols, p, salt, _ptr(out)))
    return out[: rows * cols]


def cyclic_apply_cpu(desc, dr, dc, rounds: int, pairs, x, width: int) -> np.ndarray:
    """S x on the host for a cyclic operator description (pgw_cyclic_apply_cpu:
    the reference's cpu_apply_eliminated_square, matrix-free; no device)."""
    x = _u32(x).reshape(-1)
    dr, dc = _u32(dr).reshape(-1), _u32(dc).reshape(-1)
    pr = _u32(pairs).reshape(-1) if pairs is not None and len(pairs) else np.zeros(4, np.uint32)
    out = np.empty(max(1, x.size), np.uint32)
    _check(lib().pgw_cyclic_apply_cpu(ctypes.byref(desc), _ptr(dr), _ptr(dc), rounds, _ptr(pr),
                                      _ptr(x), width, _ptr(out)))
    return out[: x.size]


def rowmix_realized_seed(seed: int, op_name: str, sector: int, rows: int, p: int) -> int:
    return lib().pgw_rowmix_realized_seed(seed, op_name.encode(), sector, rows, p)


def rowmix_plan(realized_seed: int, rows: int, rounds: int, p: int) -> np.ndarray:
    """rounds*(rows//2) x 4 array of (first, second, a, b)."""
    pairs = np.zeros((max(1, rounds * (rows // 2)), 4), np.uint32)
    _check(lib().pgw_rowmix_plan(realized_seed, rows, rounds, p, _ptr(pairs)))
    return pairs[: rounds * (rows // 2)]


def rowmix_plan_hash(rows: int, rounds: int, seed: int, realized: int, pairs: np.ndarray) -> int:
    pairs = _u32(pairs) if len(pairs) else np.zeros((1, 4), np.uint32)
    return lib().pgw_rowmix_plan_hash(rows, rounds, seed, realized, _ptr(pairs))


def fnv1a64_u32(values) -> int:
    v = _u32(values)
    return lib().pgw_fnv1a64_u32(_ptr(v) if v.size else None, v.size)


def synthetic_table(seed: int, p: int, count: int) -> np.ndarray:
    out = np.empty(count, np.uint32)
    _check(lib().pgw_synthetic_table(seed, p, count, _ptr(out)))
    return out


def sequence_length(max_len: int) -> int:
    """L = 1 + 4*ceil((max_len-1)/4) (cuprym_cyclic.cu:2425)."""
    return 1 + 4 * ((max_len - 1 + 3) // 4)


# ---------------------------------------------------------------------------
# specs
# ---------------------------------------------------------------------------

@dataclass
class SyntheticSpec:
    """A = E o D on Lambda^k V (x) R_d with tables drawn from two splitmix64 seeds."""
    n: int
    k: int
    m_src: int
    m_tgt: int
    prime: int
    table_seed_d: int
    table_seed_e: int

    def tables(self):
        count = self.n * self.m_tgt * self.m_src
        return (synthetic_table(self.table_seed_d, self.prime, count),
                synthetic_table(self.table_seed_e, self.prime, count))

    @property
    def cols(self) -> int:
        from math import comb
        return comb(self.n, self.k) * self.m_src


def synthetic_spec_from_json(obj: dict) -> SyntheticSpec:
    s = obj["synthetic"]
    return SyntheticSpec(int(s["n"]), int(s["k"]), int(s["M"]), int(s["M_prime"]),
                         int(obj["modulus"]), int(s["table_seed_d"]), int(s["table_seed_e"]))


def load_cyclic_instance(path) -> dict:
    with open(path) as f:
        return json.load(f)


# ---------------------------------------------------------------------------
# context and operator
# ---------------------------------------------------------------------------

class Context:
    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().pgw_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    def synchronize(self):
        _check(lib().pgw_ctx_synchronize(self.h))

    def device_info(self):
        name = ctypes.create_string_buffer(256)
        total = ctypes.c_uint64()
        _check(lib().pgw_ctx_device_info(self.h, name, 256, ctypes.byref(total)))
        return name.value.decode(), total.value

    def gram_upper_noalloc(self, prime: int, block: int, rows: int, d_x: int, d_y: int,
                           d_out: int, stream: int | None = None):
        """Packed upper triangle of X^T Y mod p (pgw_gram_upper_noalloc): no
        operator and no device allocation; calls on one GPU must not overlap."""
        _check(lib().pgw_gram_upper_noalloc(self.h, prime, block, rows, d_x, d_y, d_out, stream))

    def close(self):
        """Free the context; refused (PgwError) while operators, batches or
        servers created on it are still open."""
        if getattr(self, "h", None):
            _check(lib().pgw_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Operator:
    """One sector operator on the GPU (pgw_operator_create)."""

    def __init__(self, ctx: Context, desc: OperatorDesc, dr, dc, rounds: int, pairs, keep):
        self.ctx = ctx
        dr, dc = _u32(dr), _u32(dc)
        pairs_arr = _u32(pairs).reshape(-1, 4) if pairs is not None and len(pairs) else None
        h = ctypes.c_void_p()
        _check(lib().pgw_operator_create(ctx.h, ctypes.byref(desc), _ptr(dr), _ptr(dc), rounds,
                                         _ptr(pairs_arr) if pairs_arr is not None else None,
                                         ctypes.byref(h)))
        self.h = h
        info = OpInfo()
        _check(lib().pgw_operator_info(self.h, ctypes.byref(info)))
        self.rows, self.cols, self.mid = info.rows, info.cols, info.mid_rows
        self.device_bytes = info.device_bytes
        self.method = info.method
        self.block = info.block
        self.prime = desc.prime

    # -- construction ------------------------------------------------------
    @staticmethod
    def synthetic_desc(spec: SyntheticSpec, block: int, method: int, table_d, table_e):
        d = OperatorDesc()
        d.kind, d.prime, d.block, d.method = KIND_SYNTHETIC, spec.prime, block, method
        d.n, d.k, d.m_src, d.m_tgt = spec.n, spec.k, spec.m_src, spec.m_tgt
        d.table_d, d.table_e = table_d.ctypes.data, table_e.ctypes.data
        return d

    @classmethod
    def synthetic(cls, ctx, spec: SyntheticSpec, block: int, dr, dc, rounds=0, pairs=None,
                  method=METHOD_AUTO, tables=None):
        td, te = tables if tables is not None else spec.tables()
        td, te = _u32(td), _u32(te)
        d = cls.synthetic_desc(spec, block, method, td, te)
        return cls(ctx, d, dr, dc, rounds, pairs, [td, te])

    @staticmethod
    def cyclic_desc(inst: dict, sector: int, block: int):
        g, p, r = int(inst["genus"]), int(inst["modulus"]), int(inst["cyclic"]["order"])
        w = inst["weights"]
        e = inst["third_section_elimination"]
        arrays = dict(
            mu=_u32(np.asarray(inst["artinian"]["mu"]["values"]).reshape(-1)),
            wv=np.ascontiguousarray(np.asarray(w["V"], np.uint16)),
            wa0=np.ascontiguousarray(np.asarray(w["A0"], np.uint16)),
            wa1=np.ascontiguousarray(np.asarray(w["A1"], np.uint16)),
            kw=np.ascontiguousarray(np.asarray(e["kernel_weights"], np.uint16)),
            R=_u32(np.asarray(e["right_inverse_A0_by_A1"]).reshape(-1)),
            B=_u32(np.asarray(e["kernel_basis_A0_by_3"]).reshape(-1)),
        )
        d = OperatorDesc()
        d.kind, d.prime, d.block, d.method = KIND_CYCLIC, p, block, METHOD_PLAIN
        d.genus, d.order, d.sector = g, r, sector
        d.mu = arrays["mu"].ctypes.data
        d.weight_v, d.weight_a0, d.weight_a1 = (arrays["wv"].ctypes.data, arrays["wa0"].ctypes.data,
                                                arrays["wa1"].ctypes.data)
        d.w_index, d.w_weight = int(e["w_V_index"]), int(e["w_weight"])
        d.kernel_dim = arrays["kw"].size
        d.kernel_weights = arrays["kw"].ctypes.data
        d.right_inverse, d.kernel_basis = arrays["R"].ctypes.data, arrays["B"].ctypes.data
        return d, arrays

    @classmethod
    def cyclic(cls, ctx, inst: dict, sector: int, block: int, dr, dc, rounds=0, pairs=None):
        d, keep = cls.cyclic_desc(inst, sector, block)
        return cls(ctx, d, dr, dc, rounds, pairs, keep)

    @staticmethod
    def deformation_desc(inst: dict, sector: int, block: int, variant: str, dense, drops=(0, 0)):
        """Augmented [F | -dense] or replacement [F without drops | dense]
        (cuda/cuprym_deformation.cu:2554-2855); dense is rows x 2."""
        d, arrays = Operator.cyclic_desc(inst, sector, block)
        d.kind = {"augmented": KIND_CYCLIC_AUGMENTED, "replacement": KIND_CYCLIC_REPLACEMENT}[variant]
        dense = np.ascontiguousarray(np.asarray(dense, dtype=np.uint32))
        if dense.ndim != 2:
            raise ValueError("dense block must be 2-dimensional (target rows x columns)")
        arrays["dense"] = dense
        d.dense_cols, d.dense_block = dense.shape[1], dense.ctypes.data
        d.drop_cols[0], d.drop_cols[1] = int(drops[0]), int(drops[1])
        return d, arrays

    @classmethod
    def deformation(cls, ctx, inst: dict, sector: int, block: int, variant: str, dense, drops,
                    dr, dc, rounds=0, pairs=None):
        d, keep = cls.deformation_desc(inst, sector, block, variant, dense, drops)
        return cls(ctx, d, dr, dc, rounds, pairs, keep)

    @staticmethod
    def dims(desc: OperatorDesc):
        rows, cols = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().pgw_operator_dims(ctypes.byref(desc), ctypes.byref(rows), ctypes.byref(cols)))
        return rows.value, cols.value

    # -- device-pointer entry points ------------------------------------------
    def apply_square(self, d_in: int, d_out: int, stream: int | None = None):
        _check(lib().pgw_apply_square(self.h, d_in, d_out, stream))

    def apply_forward(self, d_in: int, d_out: int, stream: int | None = None):
        _check(lib().pgw_apply_forward(self.h, d_in, d_out, stream))

    def gram_upper(self, d_x: int, d_y: int, d_out: int, stream: int | None = None):
        _check(lib().pgw_gram_upper(self.h, d_x, d_y, d_out, stream))

    def set_gram(self, method: int):
        """Block projections on int8 MFMA (METHOD_MFMA), the vector ALUs
        (METHOD_PLAIN) or the default (METHOD_AUTO) (pgw_operator_set_gram)."""
        _check(lib().pgw_operator_set_gram(self.h, method))

    def prepare_row_mixing(self):
        """Build the row-mixing chains once, so that every later apply runs
        P^T Dr P in one pass (pgw_operator_prepare_row_mixing); a no-op for
        operators it does not apply to."""
        _check(lib().pgw_operator_prepare_row_mixing(self.h))

    def bench_apply(self, d_a: int, d_b: int, iters: int):
        total, sweep = ctypes.c_double(), ctypes.c_double()
        count = ctypes.c_uint32()
        _check(lib().pgw_bench_apply(self.h, d_a, d_b, iters, ctypes.byref(total),
                                     ctypes.byref(sweep), ctypes.byref(count)))
        return total.value, sweep.value, count.value

    STAGES = ("sweep", "digit_split", "row_mixing", "row_scaling", "sparse_maps")

    def bench_stages(self, d_a: int, d_b: int, iters: int = 3):
        """Per-stage time of one apply (pgw_bench_stages): {stage: (ms, bytes, launches)}."""
        ms = np.zeros(len(self.STAGES), np.float64)
        by = np.zeros(len(self.STAGES), np.float64)
        n = np.zeros(len(self.STAGES), np.uint32)
        _check(lib().pgw_bench_stages(self.h, d_a, d_b, iters, _ptr(ms), _ptr(by), _ptr(n)))
        return {s: (float(ms[q]), float(by[q]), int(n[q])) for q, s in enumerate(self.STAGES) if n[q]}

    def autotune(self, reps: int = 3):
        """Pick the fastest MFMA tile configuration per sweep; returns (choices, ms)."""
        choice = np.zeros(4, np.uint32)
        ms = np.zeros(4, np.float64)
        _check(lib().pgw_operator_autotune(self.h, reps, _ptr(choice), _ptr(ms)))
        return choice.tolist(), ms.tolist()

    # -- host-buffer entry points ------------------------------------------------
    def sequence(self, v, max_len: int, progress=None):
        v = _u32(v).reshape(-1)
        b = self.block
        ut = b * (b + 1) // 2
        L = sequence_length(max_len)
        seq = np.empty(L * ut, np.uint32)
        curv = np.empty(self.cols * b, np.uint32)
        out_len = ctypes.c_uint64()
        cb = PROGRESS_FN(lambda done, total, user: progress(done, total)) if progress else PROGRESS_FN()
        _check(lib().pgw_sequence_run(self.h, _ptr(v), max_len, _ptr(seq), _ptr(curv),
                                      ctypes.byref(out_len), cb, None))
        assert out_len.value == L
        return seq.reshape(L, ut), curv

    def kernel_vectors(self, v, gens):
        v = _u32(v).reshape(-1)
        gens = _u32(gens).reshape(-1)
        b = self.block
        ngen = gens.size // (b * b)
        d = np.empty(self.cols * b, np.uint32)
        prod = np.empty(self.rows * b, np.uint32)
        _check(lib().pgw_kernel_vectors(self.h, _ptr(v), _ptr(gens), ngen, _ptr(d), _ptr(prod)))
        return d, prod

    def close(self):
        if getattr(self, "h", None):
            lib().pgw_operator_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many sector sequence loops advanced together (pgw_batch_*): one round
    of every job per replay on one stream; compatible MFMA jobs share one
    launch per stage (grouped), a round replays as a captured graph."""

    GRAPH, GROUPED, AUTOTUNE, PERSISTENT = 1, 2, 4, 8

    def __init__(self, ctx: Context, use_graph: bool = True, grouped: bool = True,
                 autotune: bool = False, persistent: bool = False):
        self.ctx = ctx
        h = ctypes.c_void_p()
        flags = ((self.GRAPH if use_graph else 0) | (self.GROUPED if grouped else 0)
                 | (self.AUTOTUNE if autotune else 0) | (self.PERSISTENT if persistent else 0))
        _check(lib().pgw_batch_create(ctx.h, flags, ctypes.byref(h)))
        self.h = h
        self.jobs = {}

    def groups(self):
        """(groups, grouped jobs) of the current plan."""
        g, n = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().pgw_batch_groups(self.h, ctypes.byref(g), ctypes.byref(n)))
        return g.value, n.value

    def add(self, op: Operator, v, max_len: int) -> int:
        v = _u32(v).reshape(-1)
        job = ctypes.c_uint64()
        _check(lib().pgw_batch_add(self.h, op.h, _ptr(v), max_len, ctypes.byref(job)))
        self.jobs[job.value] = (op, max_len)
        return job.value

    def step(self, max_replays: int = 1) -> int:
        """Up to max_replays replays; returns the number of unfinished jobs."""
        active = ctypes.c_uint64()
        _check(lib().pgw_batch_step(self.h, max_replays, ctypes.byref(active)))
        return active.value

    def done(self, job: int) -> bool:
        d = ctypes.c_int()
        _check(lib().pgw_batch_poll(self.h, job, ctypes.byref(d)))
        return bool(d.value)

    def result(self, job: int):
        op, max_len = self.jobs.pop(job)
        b = op.block
        ut = b * (b + 1) // 2
        L = sequence_length(max_len)
        seq = np.empty(L * ut, np.uint32)
        curv = np.empty(op.cols * b, np.uint32)
        n = ctypes.c_uint64()
        _check(lib().pgw_batch_result(self.h, job, _ptr(seq), _ptr(curv), ctypes.byref(n)))
        assert n.value == L
        return seq.reshape(L, ut), curv

    def run(self, max_replays: int = 0):
        """Step until every job is done; returns {job: (seq, curv)}."""
        while self.step(max_replays or (1 << 30)):
            pass
        return {j: self.result(j) for j in list(self.jobs)}

    def terms(self, job: int, first: int = 0, seq=None) -> int:
        """Terms of a running job that are already final (pgw_batch_terms);
        with `seq` (a uint32 (L, b(b+1)/2) array) rows [first, done) are
        copied into it."""
        done = ctypes.c_uint64()
        _check(lib().pgw_batch_terms(self.h, job, first, None if seq is None else _ptr(seq),
                                     ctypes.byref(done)))
        return done.value

    def graph_builds(self) -> int:
        n = ctypes.c_uint64()
        _check(lib().pgw_batch_graph_builds(self.h, ctypes.byref(n)))
        return n.value

    def close(self):
        if getattr(self, "h", None):
            lib().pgw_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_wdm(path, rows: int, cols: int, p: int, dr, dc, v, curv, b: int, seq, threads: int = 0):
    """Symmetric WDM file (one zstd frame for .zst paths) via pgw_wdm_write;
    seq is (L, b(b+1)/2), v and curv (cols, b)."""
    dr, dc = _u32(dr).reshape(-1), _u32(dc).reshape(-1)
    v, curv, seq = _u32(v).reshape(-1), _u32(curv).reshape(-1), _u32(seq)
    L = seq.shape[0] if seq.ndim == 2 else seq.size // (b * (b + 1) // 2)
    if dr.size != rows or dc.size != cols or v.size != cols * b or curv.size != cols * b or \
            seq.size != L * (b * (b + 1) // 2):
        raise ValueError("WDM array shapes do not match rows, cols and b")
    _check(lib().pgw_wdm_write(str(path).encode(), rows, cols, p, _ptr(dr), _ptr(dc), _ptr(v),
                               _ptr(curv), b, _ptr(seq.reshape(-1)), L, threads))


class WdmStream:
    """The same WDM file as write_wdm, written while the sequence is still
    being computed (pgw_wdm_stream_*): the prefix (header, Dr, Dc, V) at once,
    sequence rows as terms(seq, done) hands them over, curV and the entry
    lines at finish(curv).  `seq` must be the (L, b(b+1)/2) array whose rows
    [0, done) are final at each terms() call."""

    def __init__(self, path, rows: int, cols: int, p: int, dr, dc, v, b: int, length: int,
                 threads: int = 0, max_text_bytes: int = 1 << 31):
        self._keep = (_u32(dr).reshape(-1), _u32(dc).reshape(-1), _u32(v).reshape(-1))
        dr, dc, v = self._keep
        if dr.size != rows or dc.size != cols or v.size != cols * b:
            raise ValueError("WDM array shapes do not match rows, cols and b")
        self.length, self.ut, self.cols, self.b = length, b * (b + 1) // 2, cols, b
        self.h = ctypes.c_void_p()
        self._seq = None
        _check(lib().pgw_wdm_stream_open(str(path).encode(), rows, cols, p, _ptr(dr), _ptr(dc), _ptr(v),
                                         b, length, threads, max_text_bytes, ctypes.byref(self.h)))

    def terms(self, seq, done: int):
        if self._seq is None:
            if seq.dtype != np.uint32 or not seq.flags.c_contiguous or seq.size != self.length * self.ut:
                raise ValueError("seq must be a contiguous uint32 (L, b(b+1)/2) array")
            self._seq = seq  # the writer reads it until finish
        elif seq is not self._seq:
            raise ValueError("terms() must be given the same array every time")
        _check(lib().pgw_wdm_stream_terms(self.h, _ptr(seq), done))

    def finish(self, curv):
        curv = _u32(curv).reshape(-1)
        if curv.size != self.cols * self.b or self._seq is None:
            raise ValueError("finish needs the sequence and a (cols, b) curV block")
        _check(lib().pgw_wdm_stream_finish(self.h, _ptr(curv)))
        self.close()

    def close(self):
        if self.h:
            lib().pgw_wdm_stream_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def job_footprint(desc: OperatorDesc, max_len: int, rounds: int = 2) -> int:
    """Device bytes a sector job is admitted with (pgw_job_footprint)."""
    out = ctypes.c_uint64()
    _check(lib().pgw_job_footprint(ctypes.byref(desc), max_len, rounds, ctypes.byref(out)))
    return out.value


class Server:
    """Sector-job serving loop with HBM-footprint admission (pgw_server_*)."""

    QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3

    def __init__(self, ctx: Context, hbm_budget: int, max_active: int = 16,
                 replays_per_step: int = 1):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(lib().pgw_server_create(ctx.h, hbm_budget, max_active, replays_per_step,
                                       ctypes.byref(h)))
        self.h = h
        self.meta = {}

    def submit(self, desc: OperatorDesc, dr, dc, v, max_len: int, rounds: int = 0,
               pairs=None, keep=None) -> int:
        """Queue a job; desc's arrays (kept alive by `keep`) are copied at once."""
        dr, dc, v = _u32(dr), _u32(dc), _u32(v).reshape(-1)
        pairs_arr = _u32(pairs).reshape(-1, 4) if pairs is not None and len(pairs) else None
        t = ctypes.c_uint64()
        _check(lib().pgw_server_submit(self.h, ctypes.byref(desc), _ptr(dr), _ptr(dc), rounds,
                                       _ptr(pairs_arr) if pairs_arr is not None else None,
                                       _ptr(v), max_len, ctypes.byref(t)))
        del keep
        self.meta[t.value] = (desc.block, dc.size, max_len)
        return t.value

    def wait(self, ticket: int, timeout: float = -1.0) -> int:
        st = ctypes.c_int()
        _check(lib().pgw_server_wait(self.h, ticket, timeout, ctypes.byref(st)))
        return st.value

    def result(self, ticket: int):
        """(seq, curv, times) of a finished job; times = seconds since the server
        started at: submitted, admitted, in the batch, out of the batch, finished."""
        b, cols, max_len = self.meta.pop(ticket)
        ut = b * (b + 1) // 2
        L = sequence_length(max_len)
        seq = np.empty(L * ut, np.uint32)
        curv = np.empty(cols * b, np.uint32)
        times = np.zeros(5, np.float64)
        n = ctypes.c_uint64()
        _check(lib().pgw_server_result(self.h, ticket, _ptr(seq), _ptr(curv), ctypes.byref(n),
                                       _ptr(times)))
        return seq.reshape(L, ut), curv, tuple(times.tolist())

    def enable_rank(self, threads: int = 1, stopping_threshold: int = DEFAULT_STOPPING_THRESHOLD,
                    max_nlen_buffer: int = DEFAULT_MAX_NLEN_BUFFER, keep_generator: bool = False):
        """Run the rank consumer on every job submitted afterwards (host workers)."""
        _check(lib().pgw_server_enable_rank(self.h, threads, stopping_threshold, max_nlen_buffer,
                                            1 if keep_generator else 0))
        self.keep_generator = keep_generator

    def rank(self, ticket: int):
        """(RankStep, generator (terms, b, b) or None, seconds) of a finished job;
        call before result(), which releases the job."""
        info = RankInfo()
        terms, secs = ctypes.c_uint64(), ctypes.c_double()
        _check(lib().pgw_server_rank(self.h, ticket, ctypes.byref(info), None, 0, ctypes.byref(terms),
                                     ctypes.byref(secs)))
        gen = None
        if getattr(self, "keep_generator", False):
            b = self.meta[ticket][0]
            gen = np.zeros(terms.value * b * b, np.uint32)
            _check(lib().pgw_server_rank(self.h, ticket, ctypes.byref(info), _ptr(gen), gen.size,
                                         ctypes.byref(terms), ctypes.byref(secs)))
            gen = gen.reshape(terms.value, b, b)
        return RankComputation._step(info), gen, secs.value

    def stats(self):
        c, pb, pa, q = (ctypes.c_uint64() for _ in range(4))
        _check(lib().pgw_server_stats(self.h, ctypes.byref(c), ctypes.byref(pb), ctypes.byref(pa),
                                      ctypes.byref(q)))
        return dict(completed=c.value, peak_bytes=pb.value, peak_active=pa.value, queued=q.value)

    def close(self):
        if getattr(self, "h", None):
            lib().pgw_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# rank consumer (third_party/bcw/src/rank_pipeline.rs:105-287)
# ---------------------------------------------------------------------------

@dataclass
class RankStep:
    """RankStep (rank_pipeline.rs:53-60) plus the basis' max_nlen."""
    processed_sequence_len: int
    delta_spread: int
    rank_lower_bound: int
    max_nlen: int
    status: int


class RankComputation:
    """RankComputation: shifted order basis of [S(x); I_b] over the usable
    terms of a symmetric block sequence (term 0 = V^T V is dropped).

    `step(seq)` takes the whole sequence so far, shape (L, b(b+1)/2), and
    consumes the terms after those already processed; it raises PgwError with
    status NEED_MORE_SEQUENCE when there are none."""

    def __init__(self, prime: int, block: int, n_rows: int, n_cols: int,
                 stopping_threshold: int = DEFAULT_STOPPING_THRESHOLD,
                 max_nlen_buffer: int = DEFAULT_MAX_NLEN_BUFFER, threads: int = 0):
        self.h = None
        self.block = block
        h = ctypes.c_void_p()
        _check(lib().pgw_rank_create(prime, block, n_rows, n_cols, stopping_threshold,
                                     max_nlen_buffer, threads, ctypes.byref(h)))
        self.h = h

    @staticmethod
    def _step(info: RankInfo) -> RankStep:
        return RankStep(info.processed, info.delta_spread, info.rank, info.max_terms, info.status)

    def step(self, seq) -> RankStep:
        ut = self.block * (self.block + 1) // 2
        seq = _u32(seq).reshape(-1)
        if seq.size % ut:
            raise ValueError(f"sequence size {seq.size} is not a multiple of b(b+1)/2 = {ut}")
        info = RankInfo()
        _check(lib().pgw_rank_step(self.h, _ptr(seq), seq.size // ut, ctypes.byref(info)))
        return self._step(info)

    def info(self) -> RankStep:
        info = RankInfo()
        _check(lib().pgw_rank_get_info(self.h, ctypes.byref(info)))
        return self._step(info)

    @property
    def rank_lower_bound(self) -> int:
        return self.info().rank_lower_bound

    @property
    def delta_spread(self) -> int:
        return self.info().delta_spread

    @property
    def status(self) -> int:
        return self.info().status

    def delta(self) -> np.ndarray:
        out = np.zeros(2 * self.block, np.uint64)
        _check(lib().pgw_rank_delta(self.h, _ptr(out)))
        return out

    def generator(self) -> np.ndarray:
        """The generator matrices (terms, b, b): line k of `<stem>_generators.txt`
        is matrix k, row-major."""
        b = self.block
        terms = ctypes.c_uint64()
        _check(lib().pgw_rank_generator(self.h, None, 0, ctypes.byref(terms)))
        out = np.zeros(terms.value * b * b, np.uint32)
        _check(lib().pgw_rank_generator(self.h, _ptr(out), out.size, ctypes.byref(terms)))
        return out.reshape(terms.value, b, b)

    def close(self):
        if getattr(self, "h", None):
            lib().pgw_rank_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""The command-line drivers: option handling (CPU) and end-to-end WDM files
and kernel-vector recovery against the oracle (GPU)."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_ctypes as oc

ROOT = oc.ROOT
BIN = ROOT / "prym_green_ancillary_amd" / "bin"
GEN = BIN / "cuprym_cyclic"
REC = BIN / "cuprym_cyclic_kernel_vectors"
G12 = ROOT / "tests" / "golden" / "g12.json"


def run(args, **kw):
    return subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=600, **kw)


def decompress(path: Path) -> str:
    import ctypes
    raw = path.read_bytes()
    if raw[:4] != b"\x28\xb5\x2f\xfd":
        return raw.decode()
    # streamed frames carry no content size: use libzstd's streaming decoder
    z = ctypes.CDLL("libzstd.so.1")

    class InBuf(ctypes.Structure):
        _fields_ = [("src", ctypes.c_void_p), ("size", ctypes.c_size_t), ("pos", ctypes.c_size_t)]

    class OutBuf(ctypes.Structure):
        _fields_ = [("dst", ctypes.c_void_p), ("size", ctypes.c_size_t), ("pos", ctypes.c_size_t)]

    z.ZSTD_createDStream.restype = ctypes.c_void_p
    z.ZSTD_initDStream.argtypes = [ctypes.c_void_p]
    z.ZSTD_decompressStream.restype = ctypes.c_size_t
    z.ZSTD_decompressStream.argtypes = [ctypes.c_void_p, ctypes.POINTER(OutBuf), ctypes.POINTER(InBuf)]
    z.ZSTD_isError.argtypes = [ctypes.c_size_t]
    z.ZSTD_freeDStream.argtypes = [ctypes.c_void_p]
    ds = z.ZSTD_createDStream()
    z.ZSTD_initDStream(ds)
    src = ctypes.create_string_buffer(raw, len(raw))
    inb = InBuf(ctypes.cast(src, ctypes.c_void_p), len(raw), 0)
    chunk = ctypes.create_string_buffer(1 << 20)
    parts = []
    while True:
        outb = OutBuf(ctypes.cast(chunk, ctypes.c_void_p), len(chunk), 0)
        rc = z.ZSTD_decompressStream(ds, ctypes.byref(outb), ctypes.byref(inb))
        assert not z.ZSTD_isError(rc), "zstd error"
        parts.append(chunk.raw[:outb.pos])
        if inb.pos >= inb.size and outb.pos == 0:
            break
    z.ZSTD_freeDStream(ds)
    return b"".join(parts).decode()


@pytest.mark.parametrize("args,msg", [
    (["--sector", "9"], "--sector must be in 0..5"),
    (["--sector", "0", "--method", "fast"], "--method must be"),
    (["--sector", "0", "-v", "17"], "-v must be in 1..16"),
    (["--sector", "0", "--bogus", "1"], "unknown option"),
    (["--sector", "0", "--operator", "synthetic"], "--operator must be one of: eliminated, augmented, replacement"),
])
def test_generator_cli_rejects_bad_options(args, msg):
    assert GEN.exists(), "build the drivers first (make -C prym_green_ancillary_amd/csrc)"
    r = run([GEN, G12, *args])
    assert r.returncode == 1
    assert msg in r.stderr


def test_generator_cli_missing_instance(tmp_path):
    r = run([GEN, tmp_path / "none.json", "--sector", "0"])
    assert r.returncode == 1 and "failed to open" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("sector,b", [(0, 4), (3, 2)])
def test_generator_cli_writes_oracle_wdm(tmp_path, g12, sector, b):
    r = run([GEN, G12, "--sector", sector, "--out", tmp_path, "-v", b])
    assert r.returncode == 0, r.stderr
    name = f"eliminated-sector{sector:02d}-v{b}-rowmix-r2-s10001"
    wdm = tmp_path / f"{name}.wdm.zst"
    side = json.loads((tmp_path / f"{name}.rowmix.json").read_text())
    text = decompress(wdm)
    # oracle run of the same recipe
    p = g12["modulus"]
    o = oc.cyclic_op(g12, sector)
    dr, dc, v = oc.reference_random_state(1, sector, p, o.rows, o.cols, b)
    realized, pairs, plan_hash = oc.rowmix_plan(10001, "eliminated", sector, o.rows, p, 2)
    o.set_preconditioners(dr, dc, 2, pairs)
    max_len = 2 * o.cols // b + 50
    seq, curv = o.sequence(v, b, max_len)
    ref = tmp_path / "oracle.wdm"
    oc.lib().orc_write_wdm(str(ref).encode(), o.rows, o.cols, p, dr, dc, v, curv, b,
                           np.ascontiguousarray(seq.reshape(-1)), seq.shape[0])
    assert text == ref.read_text()
    assert side["rowmix"]["realized_seed"] == realized
    assert int(side["rowmix"]["hash_fnv1a64"], 16) == plan_hash
    assert int(side["hashes"]["initial_block_fnv1a64"], 16) == oc.lib().orc_fnv1a64_u32_vector(v, v.size)
    # refuses to overwrite
    r2 = run([GEN, G12, "--sector", sector, "--out", tmp_path, "-v", b])
    assert r2.returncode == 1 and "refusing" in r2.stderr


@pytest.mark.gpu
def test_recovery_cli_certificate(tmp_path, g12, g12_reference):
    """A WDM whose initial block is Dc^-1 K0 and the generator G_0 = I give D = Dc^-1 K0,
    a certified kernel block of rank 2; a wrong generator fails the certificate."""
    p, b, sector = g12["modulus"], 2, 0
    o = oc.cyclic_op(g12, sector)
    dr, dc, _ = oc.reference_random_state(1, sector, p, o.rows, o.cols, b)
    K0 = np.asarray(g12_reference["certificate"]["K0"], dtype=np.int64)
    inv = np.array([pow(int(c), p - 2, p) for c in dc], dtype=np.int64)
    V = ((K0 * inv[:, None]) % p).astype(np.uint32)
    wdm = tmp_path / "k.wdm"
    seq = np.zeros(3, np.uint32)
    oc.lib().orc_write_wdm(str(wdm).encode(), o.rows, o.cols, p, dr, dc, V.reshape(-1), V.reshape(-1),
                           b, seq, 1)
    gen = tmp_path / "k_generators.txt"
    gen.write_text("1 0 0 1\n")
    r = run([REC, G12, "-f", wdm, "--sector", sector, "--rowmix-rounds", "2"])
    assert r.returncode == 0, r.stderr
    rep = json.loads((tmp_path / "k.wdm_nullvectors_report.json").read_text())
    assert rep["candidate_rank"] == 2 and rep["product_nonzeros"] == [0, 0]
    lines = (tmp_path / "k.wdm_nullvectors_2.txt").read_text().splitlines()
    got = np.array([list(map(int, ln.split())) for ln in lines]).T
    assert (got == K0).all()
    # an initial block outside the kernel fails the certificate
    bad = tmp_path / "bad.wdm"
    W = np.random.default_rng(0).integers(1, p, size=V.shape).astype(np.uint32)
    oc.lib().orc_write_wdm(str(bad).encode(), o.rows, o.cols, p, dr, dc, W.reshape(-1), W.reshape(-1),
                           b, seq, 1)
    (tmp_path / "bad_generators.txt").write_text("1 0 0 1\n")
    r = run([REC, G12, "-f", bad, "--sector", sector, "--rowmix-rounds", "2"])
    assert r.returncode == 1 and "certificate failed" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("instance,sector,b", [("g12", 0, 4), ("g12", 4, 3), ("g20", 0, 4)])
def test_generator_cli_validate_rowmix_compares_device_and_host_exactly(tmp_path, instance, sector, b):
    """--validate-rowmix (validate_eliminated_cpu_cuda, cuda/cuprym_cyclic.cu:
    1970-2132): the device S X equals the exact host apply on the salted
    blocks 0x404, 0x505, 0x606, entry by entry."""
    print(f"validate-rowmix {instance} sector {sector} b={b}", flush=True)
    r = run([GEN, ROOT / "tests" / "golden" / f"{instance}.json", "--sector", sector, "-v", b,
             "--validate-rowmix", "--no-save", "--length", 9, "--out", tmp_path])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Validation matched: eliminated target rowmix transpose identity" in r.stdout
    for idx in range(3):
        assert f"Validation matched: eliminated square block {idx} (" in r.stdout, r.stdout[-2000:]
    assert "Validation matched: <S u, w> = <u, S w>" in r.stdout

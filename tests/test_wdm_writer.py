"""The WDM writer behind the C ABI (pgw_wdm_write): the text of every thread
count equals the oracle writer's (tests of orc_write_wdm pin it against the
BCW CLI fixture), and a .zst file is exactly one zstd frame, which the
reference's reader requires (cuda/zstd_compat.h:150-156 refuses trailing
bytes after the frame)."""
import ctypes

import numpy as np
import pytest

import oracle_ctypes as oc
import prym_green_ancillary_amd as pg


def one_frame_text(path):
    """Decode the first zstd frame and check that nothing follows it."""
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
    raw = path.read_bytes()
    assert raw[:4] == b"\x28\xb5\x2f\xfd"
    ds = z.ZSTD_createDStream()
    z.ZSTD_initDStream(ds)
    src = ctypes.create_string_buffer(raw, len(raw))
    inb = InBuf(ctypes.cast(src, ctypes.c_void_p), len(raw), 0)
    chunk = ctypes.create_string_buffer(1 << 20)
    parts, done = [], False
    while not done:
        outb = OutBuf(ctypes.cast(chunk, ctypes.c_void_p), len(chunk), 0)
        rc = z.ZSTD_decompressStream(ds, ctypes.byref(outb), ctypes.byref(inb))
        assert not z.ZSTD_isError(rc)
        parts.append(chunk.raw[:outb.pos])
        done = rc == 0  # end of the frame
        assert done or inb.pos < inb.size or outb.pos > 0
    z.ZSTD_freeDStream(ds)
    assert inb.pos == inb.size, "bytes after the first zstd frame"
    return b"".join(parts).decode()


def sample(rows, cols, L, b, p, seed):
    rng = np.random.default_rng(seed)
    dr = rng.integers(1, p, rows, dtype=np.uint32)
    dc = rng.integers(1, p, cols, dtype=np.uint32)
    v = rng.integers(0, p, (cols, b), dtype=np.uint32)
    curv = rng.integers(0, p, (cols, b), dtype=np.uint32)
    seq = rng.integers(0, p, (L, b * (b + 1) // 2), dtype=np.uint32)
    return dr, dc, v, curv, seq


@pytest.mark.parametrize("rows,cols,L,b,p", [
    (7, 9, 5, 1, 109),
    # lines longer than one formatting piece (2^18 values): split over threads
    (300_001, 600_003, 270_000, 2, 2147483647),
])
def test_writer_text_matches_oracle_for_every_thread_count(tmp_path, rows, cols, L, b, p):
    dr, dc, v, curv, seq = sample(rows, cols, L, b, p, rows)
    ref = tmp_path / "oracle.wdm"
    oc.lib().orc_write_wdm(str(ref).encode(), rows, cols, p, dr, dc, v.reshape(-1), curv.reshape(-1), b,
                           np.ascontiguousarray(seq.reshape(-1)), L)
    want = ref.read_text()
    for threads in (1, 3, 8):
        plain = tmp_path / f"t{threads}.wdm"
        pg.write_wdm(plain, rows, cols, p, dr, dc, v, curv, b, seq, threads)
        assert plain.read_text() == want, threads
        packed = tmp_path / f"t{threads}.wdm.zst"
        pg.write_wdm(packed, rows, cols, p, dr, dc, v, curv, b, seq, threads)
        assert one_frame_text(packed) == want, threads
        assert not (tmp_path / f"t{threads}.wdm.zst.tmp").exists()


def test_writer_errors(tmp_path):
    dr, dc, v, curv, seq = sample(3, 4, 2, 1, 109, 0)
    with pytest.raises(pg.PgwError, match="failed to open"):
        pg.write_wdm(tmp_path / "missing" / "x.wdm", 3, 4, 109, dr, dc, v, curv, 1, seq)
    with pytest.raises(ValueError):
        pg.write_wdm(tmp_path / "x.wdm", 3, 5, 109, dr, dc, v, curv, 1, seq)


@pytest.mark.parametrize("rows,cols,L,b,p,cap", [
    (11, 13, 41, 4, 661, 1 << 31),           # entry lines formatted as the terms arrive
    (5, 6, 9, 1, 109, 1 << 31),
    (11, 13, 41, 4, 661, 0),                 # over the text cap: formatted at finish
    (2_000, 3_000, 20_001, 16, 2147483647, 1 << 31),
])
def test_streamed_writer_text_equals_writer(tmp_path, rows, cols, L, b, p, cap):
    # the sequence arrives in uneven slices while the prefix is being written,
    # as pgw_sequence_run's progress callback hands it over; the bytes must be
    # those of the one-shot writer (itself pinned to the oracle above)
    dr, dc, v, curv, seq = sample(rows, cols, L, b, p, L)
    want_path = tmp_path / "whole.wdm"
    pg.write_wdm(want_path, rows, cols, p, dr, dc, v, curv, b, seq)
    want = want_path.read_text()
    rng = np.random.default_rng(L)
    for name in ("streamed.wdm", "streamed.wdm.zst"):
        out = tmp_path / name
        live = np.zeros_like(seq)  # rows become final only when handed over
        s = pg.WdmStream(out, rows, cols, p, dr, dc, v, b, L, threads=3, max_text_bytes=cap)
        done = 0
        while done < L:
            step = int(rng.integers(1, max(2, L // 5)))
            live[done:done + step] = seq[done:done + step]
            done = min(L, done + step)
            s.terms(live, done)
        assert not out.exists()
        s.finish(curv)
        text = one_frame_text(out) if name.endswith(".zst") else out.read_text()
        assert text == want, name
        assert not (tmp_path / (name + ".tmp")).exists()


def test_streamed_writer_abandoned_leaves_no_file(tmp_path):
    dr, dc, v, curv, seq = sample(3, 4, 9, 2, 109, 1)
    out = tmp_path / "x.wdm.zst"
    s = pg.WdmStream(out, 3, 4, 109, dr, dc, v, 2, 9)
    s.terms(seq, 5)
    s.close()  # the loop failed: no partial output survives
    assert not out.exists() and not (tmp_path / "x.wdm.zst.tmp").exists()
    with pytest.raises(pg.PgwError, match="failed to open"):
        bad = pg.WdmStream(tmp_path / "missing" / "x.wdm", 3, 4, 109, dr, dc, v, 2, 9)
        bad.terms(seq, 9)
        bad.finish(curv)

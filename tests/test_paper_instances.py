"""The reference's own paper instances (data/g20.json, g22.json, g24.json, g26.json, g28.json).

CPU: the host layout of every eliminated sector reproduces the sector shapes
that data/cases.json records (and the elimination checks mu_w R = I,
mu_w B = 0 pass while it is built).

GPU: the generator CLI runs every sector of genus 20 and genus 22, and
sectors 0 and 1 of genus 24, with the paper recipe (b = 4, seed 1, two
row-mixing rounds from seed 10001, length 2N/4 + 50) and prym-rank reads each
WDM file; every rank equals cases.json's expected_base_ranks.  The median sequence-loop time per genus
is printed beside the published one (other hardware: context only).
"""
import json
import re
import statistics
import subprocess
import sys
import time
from pathlib import Path

import pytest

import prym_green_ancillary_amd as pg

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
BIN = ROOT / "prym_green_ancillary_amd" / "bin"


def _case(genus):
    cases = json.loads((GOLDEN / "cases.json").read_text())["cases"]
    return next(c for c in cases if c["genus"] == genus)


def _instance(genus):
    return json.loads((GOLDEN / f"g{genus}.json").read_text())


@pytest.mark.parametrize("genus,sectors", [(20, None), (22, None), (24, [0, 1]), (26, [0]), (28, [0, 1])])
def test_host_layout_matches_recorded_sector_shapes(genus, sectors):
    inst, case = _instance(genus), _case(genus)
    assert inst["modulus"] == case["prime"] and inst["cyclic"]["order"] == case["order"]
    for s in sectors if sectors is not None else range(case["order"]):
        d, keep = pg.Operator.cyclic_desc(inst, s, 4)
        assert pg.Operator.dims(d) == (case["sector_rows"][s], case["sector_columns"][s]), s


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run_genus(tmp_path, genus, sectors):
    """Generator CLI per sector, prym-rank per WDM file (ranks overlap the next
    sector's GPU run); returns {sector: (rank, loop seconds)}."""
    case = _case(genus)
    inst_path = GOLDEN / f"g{genus}.json"
    out = tmp_path / f"g{genus}"
    out.mkdir()
    ranks, loops, pending = {}, {}, []
    for s in sectors:
        _log(f"g{genus} sector {s}: generator ({case['sector_rows'][s]} x {case['sector_columns'][s]})")
        t0 = time.perf_counter()
        r = subprocess.run([str(BIN / "cuprym_cyclic"), str(inst_path), "--sector", str(s),
                            "--out", str(out), "-v", "4", "-d", "0"],
                           capture_output=True, text=True, timeout=900)
        assert r.returncode == 0, r.stderr[-2000:]
        m = re.search(r"Sector \d+ sequence runtime: (\d+) ms", r.stdout)
        assert m, r.stdout[-2000:]
        loops[s] = int(m.group(1)) / 1000.0
        _log(f"g{genus} sector {s}: sequence loop {loops[s]:.3f} s (process {time.perf_counter() - t0:.1f} s)")
        wdm = out / f"eliminated-sector{s:02d}-v4-rowmix-r2-s10001.wdm.zst"
        assert wdm.exists()
        pending.append((s, wdm, subprocess.Popen([str(BIN / "prym-rank"), str(wdm), "-t", "4"],
                                                 stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                                 text=True)))
    for s, wdm, proc in pending:
        so, se = proc.communicate(timeout=900)
        assert proc.returncode == 0, se[-2000:]
        res = (out / f"eliminated-sector{s:02d}-v4-rowmix-r2-s10001_result.txt").read_text()
        m = re.search(r"Matrix size: (\d+) x (\d+)\nRank: (\d+)", res)
        assert m, res
        assert (int(m.group(1)), int(m.group(2))) == (case["sector_rows"][s], case["sector_columns"][s])
        ranks[s] = int(m.group(3))
        _log(f"g{genus} sector {s}: rank {ranks[s]} (expected {case['expected_base_ranks'][s]})")
    med = statistics.median(loops.values())
    _log(f"PAPER g{genus}: median sequence loop {med:.3f} s over {len(loops)} sectors on this GPU; "
         f"published median {case['published_base_gpu_time_median_seconds']} s "
         f"({case['published_gpu_model']}, b = 4)")
    return ranks, loops


@pytest.mark.gpu
def test_g20_every_sector_rank_matches_cases(tmp_path):
    case = _case(20)
    ranks, _ = _run_genus(tmp_path, 20, range(case["order"]))
    assert [ranks[s] for s in range(case["order"])] == case["expected_base_ranks"]


@pytest.mark.gpu
def test_g22_every_sector_rank_matches_cases(tmp_path):
    case = _case(22)
    ranks, _ = _run_genus(tmp_path, 22, range(case["order"]))
    assert [ranks[s] for s in range(case["order"])] == case["expected_base_ranks"]


@pytest.mark.gpu
def test_g24_sectors_0_1_rank_matches_cases(tmp_path):
    # both row counts of g24 (323,302 and 323,344 rows; 147,015 sequence
    # terms each), the largest instance ranked end to end here
    case = _case(24)
    ranks, _ = _run_genus(tmp_path, 24, [0, 1])
    assert [ranks[0], ranks[1]] == case["expected_base_ranks"][:2]

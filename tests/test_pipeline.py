"""The sector-job orchestrator (prym_green_ancillary_amd/pipeline.py): the
behaviours tests/test_runner.py pins for the reference pipeline, on the CPU
with stand-in commands, and one end-to-end run of the real drivers on g12
(GPU)."""
import contextlib
import io
import json
import subprocess
import sys
import threading
import time

import numpy as np
import pytest

from prym_green_ancillary_amd import pipeline as pl
import oracle_ctypes as oc

G12 = oc.ROOT / "tests" / "golden" / "g12.json"


def make(tmp_path, gpus=(0,), threads=2, name="run"):
    return pl.SectorPipeline(G12, tmp_path / name, list(gpus), threads, prime=109)


def test_rank_file_must_match_shape_and_expected_rank(tmp_path):
    f = tmp_path / "r_result.txt"
    f.write_text("Matrix size: 100 x 90\nRank: 89\n")
    assert pl.read_rank(f, 100, 90, 89) == 89
    with pytest.raises(ValueError, match="expected 90"):
        pl.read_rank(f, 100, 90, 90)
    with pytest.raises(ValueError, match="dimensions"):
        pl.read_rank(f, 101, 90, 89)
    f.write_text("no rank here\n")
    with pytest.raises(ValueError, match="Malformed"):
        pl.read_rank(f, 100, 90, 89)


def test_kernel_report_with_a_residual_is_refused(tmp_path):
    job = dict(operator="eliminated", sector=0, rows=100, columns=90)
    f = tmp_path / "k.json"
    f.write_text(json.dumps(dict(job, prime=661, candidate_rank=2, product_nonzeros=[0, 0])))
    assert pl.validate_kernel_report(f, job, 661)["candidate_rank"] == 2
    f.write_text(json.dumps(dict(job, prime=661, candidate_rank=2, product_nonzeros=[0, 3])))
    with pytest.raises(ValueError, match="residual"):
        pl.validate_kernel_report(f, job, 661)
    f.write_text(json.dumps(dict(job, prime=109, candidate_rank=2, product_nonzeros=[0, 0])))
    with pytest.raises(ValueError, match="prime mismatch"):
        pl.validate_kernel_report(f, job, 661)


def test_result_paths_follow_the_rank_tool():
    w = "/x/eliminated-sector00-v4-rowmix-r2-s10001.wdm.zst"
    assert str(pl.result_path(w, "_result.txt")) == "/x/eliminated-sector00-v4-rowmix-r2-s10001_result.txt"
    assert str(pl.result_path(w, ".rowmix.json")).endswith("s10001.rowmix.json")


def test_failed_command_is_recorded_and_raises(tmp_path):
    run = make(tmp_path)
    with pytest.raises(RuntimeError, match="exited with 7"):
        run.command("failing", [sys.executable, "-c", "raise SystemExit(7)"])
    rec = json.loads((run.directory / "summary.json").read_text())["commands"][-1]
    assert rec["stage"] == "failing" and rec["exit_code"] == 7
    assert (run.directory / rec["log"]).is_file()
    assert not run.processes


def test_interrupt_while_waiting_terminates_and_reaps_the_child(tmp_path, monkeypatch):
    run = make(tmp_path)
    seen = []
    real_wait = subprocess.Popen.wait

    def wait(proc, timeout=None):
        if not seen:
            seen.append(proc)
            raise KeyboardInterrupt()
        return real_wait(proc, timeout=timeout)

    monkeypatch.setattr(subprocess.Popen, "wait", wait)
    with pytest.raises(KeyboardInterrupt):
        run.command("sleeper", [sys.executable, "-c", "import time; time.sleep(30)"])
    assert seen[0].poll() is not None  # terminated and reaped
    assert run.cancelled.is_set() and not run.processes
    with pytest.raises(RuntimeError, match="cancelled"):
        run.command("after", [sys.executable, "-c", "pass"])


def test_lease_pool_serves_priority_zero_first():
    pool = pl.LeasePool(["gpu0"])
    order, started = [], []
    hold = threading.Event()

    def user(tag, prio):
        started.append(tag)
        with pool.lease(1, prio):
            order.append(tag)
            if tag == "first":
                hold.wait(5)

    t0 = threading.Thread(target=user, args=("first", 1))
    t0.start()
    while not order:
        pass
    late_low = threading.Thread(target=user, args=("low", 1))
    late_low.start()
    while "low" not in started:
        pass
    time.sleep(0.05)
    high = threading.Thread(target=user, args=("high", 0))
    high.start()
    while "high" not in started:
        pass
    time.sleep(0.05)
    hold.set()
    for t in (t0, late_low, high):
        t.join(5)
    assert order == ["first", "high", "low"]
    assert pool.available == ["gpu0"]


def test_lease_pool_large_request_is_not_overtaken_and_cancel_leaves_the_queue():
    pool = pl.LeasePool(range(4))
    held = pool.acquire(3)
    got = []
    big = threading.Thread(target=lambda: got.append(pool.acquire(2)))
    big.start()
    time.sleep(0.05)
    # a later one-unit request of the same priority waits behind the two-unit one
    small = threading.Thread(target=lambda: got.append(pool.acquire(1)))
    small.start()
    time.sleep(0.05)
    assert got == []
    pool.release(held)
    big.join(5)
    small.join(5)
    assert sorted(len(g) for g in got) == [1, 2]
    # an abandoned wait (cancelled campaign) leaves no request behind
    cancelled = threading.Event()
    cancelled.set()
    with pytest.raises(RuntimeError, match="cancelled"):
        pool.acquire(4, 1, cancelled)
    for g in got:
        pool.release(g)
    assert sorted(pool.available) == [0, 1, 2, 3]
    assert sorted(pool.acquire(4)) == [0, 1, 2, 3]


def _fake_jobs(run, n):
    return [run.job("eliminated", c, 10, 9, 8) for c in range(n)]


def test_gpu_sequence_runs_overlap_bounded_cpu_rank_runs(tmp_path, monkeypatch):
    run = make(tmp_path, gpus=[0], threads=4)
    jobs = _fake_jobs(run, 3)
    by_id = {j["run_id"]: j for j in jobs}
    lock = threading.Lock()
    state = dict(gpu=0, cpu=0, peak_cpu=0, seqs=0, ranks=0, overlap=False)
    first_rank, second_rank = threading.Event(), threading.Event()

    def command(name, args, threads=None):
        job = by_id[name.rsplit("-", 1)[0]]
        wdm = pl.Path(job["wdm_path"])
        if name.endswith("-sequence"):
            with lock:
                state["gpu"] += 1
                assert state["gpu"] == 1  # one device: sequence runs never overlap
                state["seqs"] += 1
                first = state["seqs"] == 1
            try:
                if not first:
                    assert first_rank.wait(5), "GPU stalled behind CPU rank work"
                    with lock:
                        state["overlap"] |= state["cpu"] > 0
                wdm.touch()
                pl.result_path(wdm, ".rowmix.json").write_text("{}")
            finally:
                with lock:
                    state["gpu"] -= 1
        else:
            with lock:
                state["ranks"] += 1
                idx = state["ranks"]
                state["cpu"] += threads
                state["peak_cpu"] = max(state["peak_cpu"], state["cpu"])
            try:
                if idx == 1:
                    first_rank.set()
                    assert second_rank.wait(5), "independent rank runs did not overlap"
                elif idx == 2:
                    second_rank.set()
                pl.result_path(wdm, "_result.txt").write_text(
                    f"Matrix size: {job['rows']} x {job['columns']}\nRank: {job['expected_rank']}\n")
            finally:
                with lock:
                    state["cpu"] -= threads

    monkeypatch.setattr(run, "command", command)
    run.execute_jobs(jobs)
    assert state["overlap"]
    assert state["peak_cpu"] == run.threads  # two rank slots x two threads each
    assert len(run.summary["ranks"]) == 3


def test_parallel_failure_terminates_the_other_child(tmp_path, monkeypatch):
    run = make(tmp_path, gpus=[0, 1])
    jobs = _fake_jobs(run, 2)
    marker = tmp_path / "second-started"
    fail_after_marker = ("import pathlib,sys,time\nend=time.monotonic()+5\n"
                         "while not pathlib.Path(sys.argv[1]).exists() and time.monotonic()<end:\n"
                         "    time.sleep(0.01)\nraise SystemExit(7)\n")
    sleep_forever = "import pathlib,sys,time; pathlib.Path(sys.argv[1]).touch(); time.sleep(30)"
    real_command = run.command
    real_popen = subprocess.Popen
    children = []

    def popen(*a, **kw):
        p = real_popen(*a, **kw)
        children.append(p)
        return p

    def command(name, args, threads=None):
        script = fail_after_marker if name == jobs[0]["run_id"] + "-sequence" else sleep_forever
        real_command(name, [sys.executable, "-c", script, marker], threads)

    monkeypatch.setattr(run, "command", command)
    monkeypatch.setattr(subprocess, "Popen", popen)
    with contextlib.redirect_stdout(io.StringIO()):
        with pytest.raises(RuntimeError, match="exited with"):
            run.execute_jobs(jobs)
    assert marker.exists() and len(children) == 2
    assert all(p.poll() is not None for p in children)
    assert run.cancelled.is_set() and not run.processes
    assert sorted(run.devices.available) == [0, 1]  # both devices released


def test_deformation_hook_starts_before_the_other_base_jobs_finish(tmp_path, monkeypatch):
    run = make(tmp_path)
    jobs = [dict(sector=c) for c in range(2)]
    background_started, hook_done = threading.Event(), threading.Event()

    def execute_job(job):
        if job["sector"] == 1:
            background_started.set()
            assert hook_done.wait(5), "deformation waited for every base sector"

    def deform(base):
        assert base is jobs
        assert background_started.wait(5)
        hook_done.set()

    monkeypatch.setattr(run, "execute_job", execute_job)
    run.execute_jobs(jobs, deform=deform)
    assert hook_done.is_set()


@pytest.mark.gpu
def test_g12_pipeline_end_to_end(tmp_path, g12, g12_reference):
    """All six eliminated sectors (sector 0 with kernel recovery), then, from
    the deformation hook, augmented sector 1 and the replacement of sector 0,
    through the real drivers; every rank read from prym-rank's files."""
    import test_deformation as td
    p = g12["modulus"]
    run = pl.SectorPipeline(G12, tmp_path / "g12", [0], 4, prime=p)
    prof = g12_reference["sector_profiles"]
    jobs = [run.job("eliminated", s["sector"], s["rows"], s["columns"], s["rank"], kernel=s["sector"] == 0)
            for s in prof]

    def deform(base):
        print("pipeline: deformation stage", flush=True)
        dense, _ = td.augmented_block(g12, g12_reference, 1)
        f = tmp_path / "dense1.json"
        f.write_text(json.dumps({"rows": int(dense.shape[0]), "columns": 2, "prime": p, "values": dense.tolist()}))
        corr = {c["sector"]: c for c in g12_reference["first_order_corrections"]}[1]
        aug = run.job("augmented", 1, prof[1]["rows"], prof[1]["columns"] + 2,
                      corr["augmented_shape"][1] - corr["augmented_nullity"], dense=f)
        Z0, J = td.replacement_parts(g12_reference)
        zf, df = tmp_path / "z0.json", tmp_path / "drop.json"
        zf.write_text(json.dumps({"rows": int(Z0.shape[0]), "columns": 2, "prime": p, "values": Z0.tolist()}))
        df.write_text(json.dumps({"indices": J}))
        rep = run.job("replacement", 0, prof[0]["rows"], prof[0]["columns"], g12_reference["completion_rank"],
                      dense=zf, drop=df)
        run.execute_jobs([aug, rep])
        base.extend([aug, rep])

    print("pipeline: base sectors", flush=True)
    run.execute_jobs(jobs, deform=deform)
    run.finish()
    summary = json.loads((run.directory / "summary.json").read_text())
    got = {(r["operator"], r["sector"]): r["rank"] for r in summary["ranks"]}
    assert len(got) == 8
    assert got[("augmented", 1)] == 84 and got[("replacement", 0)] == 84
    assert all(got[("eliminated", s["sector"])] == s["rank"] for s in prof)
    assert all(c["exit_code"] == 0 for c in summary["commands"])

"""Sector-job orchestrator over this build's command-line drivers.

The job semantics of the reference pipeline (scripts/reproduce.py:78-352,
behaviour pinned by tests/test_runner.py) for the stages this build runs:
per sector job a GPU sequence run (``bin/cuprym_cyclic``, any operator
including the deformation variants), a CPU rank run (``bin/prym-rank``) and,
on request, GPU kernel-vector recovery (``bin/cuprym_cyclic_kernel_vectors``).

* Devices and CPU rank slots are reserved from priority queues: among waiting
  stages priority 0 runs before priority 1, arrival order breaking ties
  (Resources, scripts/reproduce.py:78-112); a device is held only while a
  GPU command runs.
* Every command logs to its own file, is recorded in summary.json (stage,
  command, exit code, seconds, log) and a nonzero exit raises
  ``RuntimeError("<stage> exited with <code>")`` (command, :142-177).
* An interrupt, or the failure of any job running in parallel, terminates
  and reaps every child process group and cancels the rest of the pipeline
  (abort, guarded_job, execute_jobs: :179-193, 282-303).
* ``execute_jobs(jobs, deform=f)`` runs ``f(jobs)`` as soon as the first job
  (sector zero) is verified, while the other sectors continue (:294-303).

The first-order and quadratic deformation stages that prepare the dense
blocks run the reference's Rust algebra and are not part of this build; the
``deform`` hook is where a caller plugs them in.
"""
from __future__ import annotations

import heapq
import itertools
import json
import os
import signal
import subprocess
import threading
import time
from concurrent.futures import ThreadPoolExecutor, as_completed
from contextlib import contextmanager
from datetime import datetime, timezone
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
BIN = PKG_DIR / "bin"

# the fixed numerical recipe of the reference's paper cases (data/cases.json settings)
DEFAULT_SETTINGS = dict(vector_count=4, numerical_seed=1, rowmix_rounds=2, rowmix_seed=10001)


def write_json(path, obj):
    tmp = Path(str(path) + ".tmp")
    tmp.write_text(json.dumps(obj, indent=2) + "\n")
    os.replace(tmp, path)


def result_path(wdm, suffix: str) -> Path:
    """<stem>_result.txt etc. next to a .wdm or .wdm.zst file (prym-rank's names)."""
    wdm = Path(wdm)
    name = wdm.name
    for ext in (".zst", ".wdm"):
        if name.endswith(ext):
            name = name[: -len(ext)]
    return wdm.with_name(name + suffix)


def read_rank(path, rows: int, columns: int, expected: int) -> int:
    """Rank of a prym-rank result file, checked against the operator's shape
    and the rank the case requires (scripts/reproduce.py:50-63)."""
    text = Path(path).read_text()
    size = rank = None
    for line in text.splitlines():
        if line.startswith("Matrix size:"):
            a, _, b = line.split(":", 1)[1].strip().partition(" x ")
            size = (int(a), int(b))
        elif line.startswith("Rank:"):
            rank = int(line.split(":", 1)[1].strip().split()[0])
    if size is None or rank is None:
        raise ValueError(f"Malformed rank file {path}")
    if size != (rows, columns):
        raise ValueError(f"Wrong matrix dimensions in {path}")
    if rank != expected:
        raise ValueError(f"Rank lower bound {rank}; expected {expected}: {path}")
    return rank


def validate_kernel_report(path, job: dict, prime: int) -> dict:
    """Kernel-vector report of a job: same operator and shape, two
    independent vectors, and F times them all zero (scripts/reproduce.py:66-76)."""
    report = json.loads(Path(path).read_text())
    want = dict(operator=job["operator"], sector=job["sector"], rows=job["rows"],
                columns=job["columns"], prime=prime)
    for key, value in want.items():
        if report.get(key) != value:
            raise ValueError(f"Kernel report {key} mismatch: {path}")
    products = report.get("product_nonzeros")
    if report.get("candidate_rank") != 2 or not products or any(x != 0 for x in products):
        raise ValueError(f"Kernel independence or residual check failed: {path}")
    return report


class Resources:
    """A pool of values (device ids, CPU slots) handed out by priority.

    A waiting reservation with a smaller priority goes first; equal
    priorities go in arrival order.  Stages already holding a value keep it
    until they finish; a reservation abandoned by an exception leaves the
    queue and wakes the others.
    """

    def __init__(self, values):
        self.available = list(values)
        self.queue = []
        self.arrivals = itertools.count()
        self.cond = threading.Condition()

    @contextmanager
    def reserve(self, priority: int = 0):
        with self.cond:
            ticket = (priority, next(self.arrivals))
            heapq.heappush(self.queue, ticket)
            try:
                self.cond.wait_for(lambda: self.available and self.queue[0] == ticket)
            except BaseException:
                self.queue.remove(ticket)
                heapq.heapify(self.queue)
                self.cond.notify_all()
                raise
            heapq.heappop(self.queue)
            value = self.available.pop(0)
            self.cond.notify_all()
        try:
            yield value
        finally:
            with self.cond:
                self.available.append(value)
                self.cond.notify_all()


class SectorPipeline:
    """Sector jobs of one instance with this build's drivers."""

    def __init__(self, instance, directory, gpus, threads: int, prime: int, settings=None, bin_dir=BIN):
        self.instance = Path(instance).resolve()
        self.prime = prime
        self.settings = dict(DEFAULT_SETTINGS, **(settings or {}))
        self.directory = Path(directory).resolve()
        self.directory.mkdir(parents=True, exist_ok=False)
        (self.directory / "logs").mkdir()
        self.bin = Path(bin_dir)
        self.gpus, self.threads = list(gpus), threads
        # a few rank workers overlap CPU rank work with the GPU sequence runs
        self.rank_slots = max(1, min(2 * len(self.gpus), threads))
        self.worker_threads = max(1, threads // self.rank_slots)
        self.devices = Resources(self.gpus)
        self.cpu_slots = Resources(range(self.rank_slots))
        self.lock = threading.Lock()
        self.processes = set()
        self.cancelled = threading.Event()
        self.summary = dict(instance=str(self.instance), prime=prime, settings=self.settings,
                            gpus=self.gpus, cpu_threads=threads, status="running",
                            started_at=datetime.now(timezone.utc).isoformat(), ranks=[], commands=[])
        self.save()

    def save(self):
        write_json(self.directory / "summary.json", self.summary)

    # -- one command ---------------------------------------------------------
    def command(self, name: str, args, threads=None):
        """Run one stage with its output in logs/<name>.log; raise on failure."""
        if self.cancelled.is_set():
            raise RuntimeError("Pipeline cancelled")
        args = [str(a) for a in args]
        log_path = self.directory / "logs" / f"{name}.log"
        env = dict(os.environ, OMP_NUM_THREADS=str(threads or self.threads))
        started = time.monotonic()
        with log_path.open("w") as log:
            with self.lock:
                if self.cancelled.is_set():
                    raise RuntimeError("Pipeline cancelled")
                proc = subprocess.Popen(args, stdout=log, stderr=subprocess.STDOUT, env=env,
                                        start_new_session=True)
                self.processes.add(proc)
            try:
                code = proc.wait()
            except BaseException:
                # an interrupt while waiting: the child stays registered until
                # abort() has terminated and reaped it
                self.abort()
                raise
            finally:
                with self.lock:
                    self.processes.discard(proc)
        record = dict(stage=name, command=args, exit_code=code, seconds=round(time.monotonic() - started, 3),
                      log=str(log_path.relative_to(self.directory)))
        with self.lock:
            self.summary["commands"].append(record)
            self.save()
        if code:
            raise RuntimeError(f"{name} exited with {code}; see {log_path}")

    def abort(self):
        """Cancel the pipeline: terminate every running child's process
        group, then reap it (SIGKILL after 5 s)."""
        self.cancelled.set()
        with self.lock:
            procs = list(self.processes)
        for proc in procs:
            try:
                os.killpg(proc.pid, signal.SIGTERM)
            except ProcessLookupError:
                pass
        for proc in procs:
            try:
                proc.wait(timeout=5)
            except subprocess.TimeoutExpired:
                try:
                    os.killpg(proc.pid, signal.SIGKILL)
                except ProcessLookupError:
                    pass
                proc.wait()

    def gpu_command(self, name, args, priority):
        """Hold a device only while the GPU command runs."""
        with self.devices.reserve(priority) as device:
            args = [*args, "--device", device]
            self.command(name, args, self.worker_threads)
            return [str(a) for a in args]

    def cpu_command(self, name, args, priority=0):
        with self.cpu_slots.reserve(priority):
            self.command(name, args, self.worker_threads)

    # -- sector jobs ---------------------------------------------------------
    def job(self, operator: str, sector: int, rows: int, columns: int, expected_rank: int,
            dense=None, drop=None, kernel=False) -> dict:
        """A job description; `columns` are those of the operator run
        (the base columns + 2 for the augmented variant)."""
        label = f"{operator}-{sector:02}"
        out = self.directory / label
        out.mkdir()
        s = self.settings
        name = (f"{operator}-sector{sector:02d}-v{s['vector_count']}"
                f"-rowmix-r{s['rowmix_rounds']}-s{s['rowmix_seed']}.wdm.zst")
        return dict(run_id=label, operator=operator, sector=sector, rows=rows, columns=columns,
                    expected_rank=expected_rank, wdm_path=str(out / name), dense=dense, drop=drop,
                    kernel=kernel)

    def execute_job(self, job: dict):
        s = self.settings
        op, sector = job["operator"], job["sector"]
        # kernel-vector and replacement jobs feed later stages: they go first
        priority = 0 if job["kernel"] or op == "replacement" else 1
        args = [self.bin / "cuprym_cyclic", self.instance, "--operator", op, "--sector", sector,
                "-v", s["vector_count"], "--seed", s["numerical_seed"],
                "--rowmix-rounds", s["rowmix_rounds"], "--rowmix-seed", s["rowmix_seed"],
                "--out", Path(job["wdm_path"]).parent]
        if job["dense"]:
            args += ["--dense-file", job["dense"]]
        if job["drop"]:
            args += ["--drop-columns", job["drop"]]
        job["args"] = self.gpu_command(job["run_id"] + "-sequence", args, priority)
        wdm = Path(job["wdm_path"])
        if not wdm.is_file() or not result_path(wdm, ".rowmix.json").is_file():
            raise RuntimeError(f"Missing sequence or row-mix metadata for {wdm}")
        args = [self.bin / "prym-rank", wdm, "--threads", self.worker_threads]
        if job["kernel"]:
            args += ["--generator"]
        self.cpu_command(job["run_id"] + "-rank", args, priority)
        job["rank"] = read_rank(result_path(wdm, "_result.txt"), job["rows"], job["columns"],
                                job["expected_rank"])
        if job["kernel"]:
            if op != "eliminated":
                raise ValueError("kernel-vector recovery runs on the eliminated operator only")
            args = [self.bin / "cuprym_cyclic_kernel_vectors", self.instance, "-f", wdm,
                    "-g", result_path(wdm, "_generators.txt"), "--operator", op, "--sector", sector]
            self.gpu_command(job["run_id"] + "-kernel", args, priority)
            validate_kernel_report(str(wdm) + "_nullvectors_report.json", job, self.prime)
        with self.lock:
            self.summary["ranks"].append({k: job[k] for k in ("operator", "sector", "rows", "columns", "rank")})
            self.save()

    def guarded_job(self, job):
        try:
            self.execute_job(job)
        except BaseException:
            self.abort()  # a failed background sector also stops the deformation lane
            raise

    def execute_jobs(self, jobs, deform=None):
        """Run jobs concurrently (resources bound the subprocesses); with
        `deform`, call deform(jobs) once jobs[0] is done, the others running on."""
        if not jobs:
            return
        with ThreadPoolExecutor(max_workers=len(jobs)) as pool:
            futures = [pool.submit(self.guarded_job, j) for j in jobs]
            try:
                if deform is not None:
                    futures[0].result()
                    deform(jobs)
                for f in as_completed(futures):
                    f.result()
            except BaseException:
                self.abort()
                raise

    def finish(self, status="ok"):
        with self.lock:
            self.summary["status"] = status
            self.summary["finished_at"] = datetime.now(timezone.utc).isoformat()
            self.save()

"""Campaign runner: the sector jobs of one instance through this build's drivers.

What a campaign does (the job semantics of the reference's scripts/
reproduce.py, whose tests/test_runner.py pins the observable behaviour): for
every sector job, a sequence run on a GPU (``bin/cuprym_cyclic``, any
operator including the deformation variants), a rank run on CPU threads
(``bin/prym-rank``) checked against the rank the case requires, and on
request kernel-vector recovery on a GPU (``bin/cuprym_cyclic_kernel_vectors``)
with its certificate checked.

How it is organised here:

* ``LeasePool`` hands out units of one resource (device ids, CPU thread
  units).  Every request gets its own wake-up event; whenever units come back
  a hand-off pass serves the waiting requests in order of priority (0 before
  1), first come first served within a priority, and stops at the first
  request that does not fit so that a large request is not starved.  A stage
  holds its lease only while its command runs.
* ``ChildRegistry`` owns the child processes.  Each child runs in a process
  group of its own; stopping it sends SIGINT to the group (the drivers then
  drop their ``.tmp`` files), escalates to SIGKILL after a grace period and
  always reaps.  ``command`` polls its child in short waits, so a cancelled
  campaign stops within a fraction of a second, and an interrupt during the
  wait stops the child before it propagates.
* ``Ledger`` keeps ``summary.json`` (every command with its exit code,
  seconds and log file, every verified rank) and rewrites it atomically after
  each change.
* ``execute_jobs`` starts one thread per job and waits on a single mailbox of
  outcomes.  The first failure cancels the campaign: the flag is raised, every
  running child is stopped, every thread is joined, then the failure is
  re-raised.  With a ``deform`` hook, the hook gets its own thread as soon as
  the first job (sector 0, the one that feeds the deformation stages) has
  succeeded, while the other sectors keep running.

The first-order and quadratic deformation stages that build the dense blocks
run the reference's Rust algebra and are outside this build; ``deform`` is
where a caller plugs them in.
"""
from __future__ import annotations

import collections
import json
import os
import queue
import signal
import subprocess
import threading
import time
from contextlib import contextmanager
from datetime import datetime, timezone
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
BIN = PKG_DIR / "bin"

# numerical recipe of the paper cases (the reference's data/cases.json settings)
DEFAULT_SETTINGS = dict(vector_count=4, numerical_seed=1, rowmix_rounds=2, rowmix_seed=10001)

POLL_S = 0.1         # child polling interval of `command`
STOP_GRACE_S = 3.0   # SIGINT -> SIGKILL escalation


def _now() -> str:
    return datetime.now(timezone.utc).isoformat()


def result_path(wdm, suffix: str) -> Path:
    """File that prym-rank (or the recovery driver) writes beside a WDM file:
    the WDM name without ``.wdm.zst`` / ``.wdm``, plus `suffix`."""
    wdm = Path(wdm)
    stem = wdm.name
    for ext in (".wdm.zst", ".wdm"):
        if stem.endswith(ext):
            stem = stem[: -len(ext)]
            break
    return wdm.with_name(stem + suffix)


def read_rank(path, rows: int, columns: int, expected: int) -> int:
    """The rank in a prym-rank result file ("Matrix size: R x C" / "Rank: r"),
    accepted only for the operator's own shape and the rank the case requires."""
    fields = {}
    for line in Path(path).read_text().splitlines():
        key, sep, value = line.partition(":")
        if sep:
            fields[key.strip()] = value.strip()
    try:
        r, c = (int(t) for t in fields["Matrix size"].split("x"))
        rank = int(fields["Rank"].split()[0])
    except (KeyError, ValueError):
        raise ValueError(f"Malformed rank result {path}: no 'Matrix size' / 'Rank' lines") from None
    if (r, c) != (rows, columns):
        raise ValueError(f"rank result {path} is for a {r} x {c} matrix, but the operator has "
                         f"dimensions {rows} x {columns}")
    if rank != expected:
        raise ValueError(f"rank result {path} gives {rank}, expected {expected}")
    return rank


def validate_kernel_report(path, job: dict, prime: int) -> dict:
    """The recovery driver's report for `job`: the same operator, sector,
    shape and prime, two independent kernel vectors and a zero residual
    F * (Dc D) in every column."""
    report = json.loads(Path(path).read_text())
    if report.get("prime") != prime:
        raise ValueError(f"kernel report {path}: prime mismatch ({report.get('prime')} != {prime})")
    for key in ("operator", "sector", "rows", "columns"):
        if report.get(key) != job[key]:
            raise ValueError(f"kernel report {path}: {key} {report.get(key)!r} is not the job's "
                             f"{job[key]!r}")
    if report.get("candidate_rank") != 2:
        raise ValueError(f"kernel report {path}: recovered vectors are not independent "
                         f"(candidate rank {report.get('candidate_rank')})")
    nonzeros = report.get("product_nonzeros") or []
    if not nonzeros or any(nonzeros):
        raise ValueError(f"kernel report {path}: nonzero residual {nonzeros}")
    return report


class LeasePool:
    """Units of one resource handed out by priority, with hand-off wake-ups."""

    def __init__(self, units):
        self.available = list(units)
        self._lock = threading.Lock()
        self._waiting = collections.defaultdict(collections.deque)  # priority -> requests

    def _hand_off(self):
        # under _lock: serve the head requests in priority order while they fit
        for prio in sorted(self._waiting):
            line = self._waiting[prio]
            while line and line[0]["count"] <= len(self.available):
                req = line.popleft()
                req["units"] = [self.available.pop(0) for _ in range(req["count"])]
                req["ready"].set()
            if line:
                return  # the head of this priority waits; lower priorities wait behind it

    def acquire(self, count: int = 1, priority: int = 1, cancelled: threading.Event | None = None):
        req = dict(count=count, units=None, ready=threading.Event())
        with self._lock:
            self._waiting[priority].append(req)
            self._hand_off()
        try:
            while not req["ready"].wait(POLL_S):
                if cancelled is not None and cancelled.is_set():
                    raise RuntimeError("campaign cancelled while waiting for a resource")
        except BaseException:
            with self._lock:
                if req["units"] is None:
                    self._waiting[priority].remove(req)
                    self._hand_off()
                    raise
            self.release(req["units"])  # granted while the wait was being abandoned
            raise
        return req["units"]

    def release(self, units):
        with self._lock:
            self.available.extend(units)
            self._hand_off()

    @contextmanager
    def lease(self, count: int = 1, priority: int = 1, cancelled=None):
        units = self.acquire(count, priority, cancelled)
        try:
            yield units
        finally:
            self.release(units)


class ChildRegistry:
    """Live child processes of a campaign, each in its own process group."""

    def __init__(self):
        self._lock = threading.Lock()
        self.live = {}

    def spawn(self, args, log, env):
        with self._lock:
            proc = subprocess.Popen(args, stdout=log, stderr=subprocess.STDOUT, env=env,
                                    start_new_session=True)
            self.live[proc.pid] = proc
        return proc

    def forget(self, proc):
        with self._lock:
            self.live.pop(proc.pid, None)

    @staticmethod
    def _signal(proc, sig):
        try:
            os.killpg(proc.pid, sig)
        except (ProcessLookupError, PermissionError):
            pass

    def stop(self, proc):
        """SIGINT to the group, SIGKILL after the grace period; always reaps."""
        if proc.poll() is None:
            self._signal(proc, signal.SIGINT)
            deadline = time.monotonic() + STOP_GRACE_S
            while proc.poll() is None and time.monotonic() < deadline:
                time.sleep(0.02)
            if proc.poll() is None:
                self._signal(proc, signal.SIGKILL)
        while True:
            try:
                proc.wait()
                break
            except KeyboardInterrupt:  # finish reaping even when interrupted again
                continue
        self.forget(proc)

    def stop_all(self):
        with self._lock:
            procs = list(self.live.values())
        for proc in procs:
            self.stop(proc)


class Ledger:
    """summary.json of a campaign, rewritten atomically after every change."""

    def __init__(self, path: Path, header: dict):
        self.path = path
        self._lock = threading.Lock()
        self.data = dict(header, status="running", started_at=_now(), commands=[], ranks=[])
        self._flush()

    def _flush(self):
        tmp = self.path.with_name(self.path.name + ".tmp")
        tmp.write_text(json.dumps(self.data, indent=2) + "\n")
        os.replace(tmp, self.path)

    def add(self, section: str, record: dict):
        with self._lock:
            self.data[section].append(record)
            self._flush()

    def close(self, status: str):
        with self._lock:
            self.data["status"] = status
            self.data["finished_at"] = _now()
            self._flush()


class SectorPipeline:
    """The sector jobs of one instance (`instance`: its JSON file) in
    `directory`, on the devices `gpus` and `threads` host threads."""

    def __init__(self, instance, directory, gpus, threads: int, prime: int, settings=None,
                 bin_dir=BIN, rank_lanes=None):
        self.instance = Path(instance).resolve()
        self.prime = prime
        self.settings = dict(DEFAULT_SETTINGS, **(settings or {}))
        self.directory = Path(directory).resolve()
        self.directory.mkdir(parents=True, exist_ok=False)
        self.logs = self.directory / "logs"
        self.logs.mkdir()
        self.bin = Path(bin_dir)
        self.gpus, self.threads = list(gpus), int(threads)
        # rank runs are CPU bound and independent: a few of them side by side
        # keep the host busy while the GPUs run the next sequences
        lanes = rank_lanes or max(1, min(2 * len(self.gpus), self.threads))
        self.rank_threads = max(1, self.threads // lanes)
        self.devices = LeasePool(self.gpus)
        self.cpu = LeasePool(range(self.threads))
        self.children = ChildRegistry()
        self.cancelled = threading.Event()
        self.ledger = Ledger(self.directory / "summary.json",
                             dict(instance=str(self.instance), prime=prime, settings=self.settings,
                                  gpus=self.gpus, cpu_threads=self.threads))

    # views the behaviour tests read
    @property
    def summary(self) -> dict:
        return self.ledger.data

    @property
    def processes(self) -> dict:
        return self.children.live

    # -- one command ---------------------------------------------------------
    def command(self, name: str, args, threads=None):
        """Run one stage with its output in logs/<name>.log and a ledger
        record; a nonzero exit raises RuntimeError("<name> exited with <code>")."""
        if self.cancelled.is_set():
            raise RuntimeError(f"{name} not started: campaign cancelled")
        args = [str(a) for a in args]
        log_path = self.logs / f"{name}.log"
        env = dict(os.environ, OMP_NUM_THREADS=str(threads or self.rank_threads))
        t0 = time.monotonic()
        with log_path.open("w") as log:
            proc = self.children.spawn(args, log, env)
            try:
                while True:
                    try:
                        code = proc.wait(timeout=POLL_S)
                        break
                    except subprocess.TimeoutExpired:
                        if self.cancelled.is_set():
                            self.children.stop(proc)
                            raise RuntimeError(f"{name} stopped: campaign cancelled") from None
            except BaseException:
                self.cancelled.set()
                self.children.stop(proc)
                raise
            self.children.forget(proc)
        self.ledger.add("commands", dict(stage=name, command=args, exit_code=code,
                                         seconds=round(time.monotonic() - t0, 3),
                                         log=str(log_path.relative_to(self.directory))))
        if code:
            raise RuntimeError(f"{name} exited with {code} (log: {log_path})")

    def _on_device(self, name, args, priority):
        with self.devices.lease(1, priority, self.cancelled) as (device,):
            args = [*args, "--device", device]
            self.command(name, args, self.rank_threads)
        return [str(a) for a in args]

    def _on_cpu(self, name, args, priority):
        with self.cpu.lease(self.rank_threads, priority, self.cancelled):
            self.command(name, args, self.rank_threads)

    # -- sector jobs ---------------------------------------------------------
    def job(self, operator: str, sector: int, rows: int, columns: int, expected_rank: int,
            dense=None, drop=None, kernel=False) -> dict:
        """Description of one sector job (`columns`: those of the operator
        run, i.e. the base columns + 2 for the augmented variant)."""
        s = self.settings
        run_id = f"{operator}-{sector:02}"
        out = self.directory / run_id
        out.mkdir()
        wdm = out / (f"{operator}-sector{sector:02d}-v{s['vector_count']}-rowmix-r{s['rowmix_rounds']}"
                     f"-s{s['rowmix_seed']}.wdm.zst")
        return dict(run_id=run_id, operator=operator, sector=sector, rows=rows, columns=columns,
                    expected_rank=expected_rank, wdm_path=str(wdm), dense=dense, drop=drop,
                    kernel=kernel)

    def execute_job(self, job: dict):
        """Sequence (GPU), rank (CPU) and, for kernel jobs, recovery (GPU)."""
        s = self.settings
        op, sector = job["operator"], job["sector"]
        if job["kernel"] and op != "eliminated":
            raise ValueError("kernel-vector recovery runs on the eliminated operator only")
        # jobs whose output feeds later stages are served first
        priority = 0 if job["kernel"] or op == "replacement" else 1
        wdm = Path(job["wdm_path"])
        gen = [self.bin / "cuprym_cyclic", self.instance, "--operator", op, "--sector", sector,
               "-v", s["vector_count"], "--seed", s["numerical_seed"],
               "--rowmix-rounds", s["rowmix_rounds"], "--rowmix-seed", s["rowmix_seed"],
               "--out", wdm.parent]
        gen += ["--dense-file", job["dense"]] if job["dense"] else []
        gen += ["--drop-columns", job["drop"]] if job["drop"] else []
        job["args"] = self._on_device(job["run_id"] + "-sequence", gen, priority)
        for produced in (wdm, result_path(wdm, ".rowmix.json")):
            if not produced.is_file():
                raise RuntimeError(f"{job['run_id']}: the sequence run left no {produced.name}")
        rank = [self.bin / "prym-rank", wdm, "--threads", self.rank_threads]
        rank += ["--generator"] if job["kernel"] else []
        self._on_cpu(job["run_id"] + "-rank", rank, priority)
        job["rank"] = read_rank(result_path(wdm, "_result.txt"), job["rows"], job["columns"],
                                job["expected_rank"])
        if job["kernel"]:
            rec = [self.bin / "cuprym_cyclic_kernel_vectors", self.instance, "-f", wdm,
                   "-g", result_path(wdm, "_generators.txt"), "--operator", op, "--sector", sector]
            self._on_device(job["run_id"] + "-kernel", rec, priority)
            validate_kernel_report(str(wdm) + "_nullvectors_report.json", job, self.prime)
        self.ledger.add("ranks", {k: job[k] for k in ("operator", "sector", "rows", "columns", "rank")})

    def cancel(self):
        """Stop the campaign: no new command starts, running children stop."""
        self.cancelled.set()
        self.children.stop_all()

    def execute_jobs(self, jobs, deform=None):
        """All jobs side by side (the resource pools bound what runs at once);
        deform(jobs), when given, starts once jobs[0] has succeeded.  The
        first failure cancels everything and is re-raised."""
        if not jobs:
            return
        mailbox = queue.Queue()

        def lane(tag, fn, *a):
            try:
                fn(*a)
                mailbox.put((tag, None))
            except BaseException as e:  # reported to the waiting thread
                mailbox.put((tag, e))

        def start(tag, fn, *a):
            t = threading.Thread(target=lane, args=(tag, fn, *a), daemon=True)
            t.start()
            return t

        threads = [start(i, self.execute_job, job) for i, job in enumerate(jobs)]
        waiting = set(range(len(jobs)))
        failure = None
        try:
            while waiting:
                tag, err = mailbox.get()
                waiting.discard(tag)
                if err is not None:
                    failure = err
                    break
                if tag == 0 and deform is not None:
                    threads.append(start("deform", deform, jobs))
                    waiting.add("deform")
        except BaseException as e:  # interrupted while waiting
            failure = e
        if failure is not None:
            self.cancel()
        for t in threads:
            while t.is_alive():
                t.join(POLL_S)
        if failure is not None:
            raise failure

    def finish(self, status: str = "ok"):
        self.ledger.close(status)

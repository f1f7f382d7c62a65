"""Block-apply throughput of one sector operator on one MI355X.

Workload (default `large`): the HBM-scale sector n=24, k=6, M=400, M'=440,
N = C(24,6)*400 = 53,838,400 rows, p = 2^31-1, 16-wide uint32 block (3.4 GB),
tables from seeds 4/5, V from seed 4, two rounds of row mixing.  One step =
one block-apply X -> S X with S = Dc F^T P^T Dr P F Dc, F = E o D, on the
int8-limb MFMA path; inputs are resident in HBM before the timed region.  The
plain u64 path is timed beside it on the same input, and the two outputs must
be bit-identical before any number is printed.

Other workloads (`--workload`): `sector` (the mid sector n=20, k=5, M=200,
M'=220, N = 3,100,800) the same way; `batch` (64 uneven sectors, batched
against one at a time); `serve` (a Poisson stream of sector jobs through the
serving loop, p50/p99 latency); `paper` (every base sector of the reference's
genus-20 or genus-22 instance, median sequence-loop time beside the published
one).

The JSON line carries `roofline` (the MFMA sweep against dense int8 peak;
`traffic` from the FETCH_SIZE/WRITE_SIZE counters committed in
profiles/traffic_<workload>.txt when they exist), `hbm_stages` (the
HBM-bound stages of the apply against 8 TB/s, timed live with HIP events) and
`cpu_baseline` (the oracle's whole apply chain, sampled, at 1 and at all
host threads).

Prints one JSON line (rank 0).  `--gpus` above 1 exits: this build runs on
one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from math import comb
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

P31 = 2147483647
# single-sector workloads: the mid sector (default) and the HBM-scale sector
WORKLOADS = {
    "sector": dict(n=20, k=5, M=200, M_prime=220, prime=P31, block=16, table_seed=0, seed=2,
                   name="single mid sector"),
    "large": dict(n=24, k=6, M=400, M_prime=440, prime=P31, block=16, table_seed=4, seed=4,
                  name="large sector at HBM scale"),
}


def macs_per_sweep(n, k, M, Mp, b):
    # every entry of every output panel (b columns) sums incidences x input-panel-rows
    # products: C(n,k-1) * (n-k+1) * M' * M * b, the same for all four sweeps
    return comb(n, k - 1) * (n - k + 1) * Mp * M * b


def cpu_baseline(spec, T, E, dr, dc, pairs, v, b, budget_s=20.0):
    """The oracle's whole apply chain (the D, E, E^T and D^T sweeps and the
    elementwise Dc / P / Dr / P^T / Dc stages), each timed on a bounded sample
    of its output rows and scaled to all rows, at one host thread and at all
    (at most 16)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_ctypes as oc

    L = oc.lib()
    nthreads = min(16, os.cpu_count() or 1)
    o = oc.synthetic_op(spec.n, spec.k, spec.m_src, spec.m_tgt, spec.prime, T, E, dr, dc, 2,
                        pairs.reshape(-1))
    N, R = spec.cols, comb(spec.n, spec.k - 1) * spec.m_tgt
    x = np.ascontiguousarray(v)
    y = np.zeros(R * b, np.uint32)          # R-space input of the E and D^T sweeps
    outs = {0: np.zeros(R * b, np.uint32), 1: np.zeros(N * b, np.uint32)}
    scratch = x.copy()

    def sampled(threads, budget):
        L.orc_set_threads(threads)
        total, notes = 0.0, []
        # growing chunks of output rows; the last (largest) chunk's time per
        # row, which keeps every thread busy, scales to the whole sweep
        for which in range(4):
            rows = L.orc_sweep_rows(o.h, which)
            xin, out = (x, outs[0]) if which in (0, 2) else (y, outs[1])
            step, done, el, per_row = max(256 * threads, rows // 4000), 0, 0.0, 0.0
            while el < budget / 5 and done < rows:
                r1 = min(rows, done + step)
                t0 = time.perf_counter()
                L.orc_sweep_partial(o.h, which, xin, out, b, done, r1)
                dt = time.perf_counter() - t0
                el, per_row = el + dt, dt / (r1 - done)
                done, step = r1, step * 2
            total += per_row * rows
            notes.append(f"{done}/{rows}")
        rows_el, el = 4096, 0.0
        while True:
            t0 = time.perf_counter()
            L.orc_elementwise_sample(o.h, scratch, b, rows_el)
            el = time.perf_counter() - t0
            if el >= budget / 10 or rows_el >= N:
                break
            rows_el = min(N, rows_el * 2)
        total += el * N / rows_el
        notes.append(f"elementwise {rows_el}/{N}")
        return total, notes

    t_single, n1 = sampled(1, budget_s * 0.4)
    t_multi, nm = sampled(nthreads, budget_s * 0.6)
    return {
        "value": 1.0 / t_multi,
        "unit": "block-applies/s",
        "cores": nthreads,
        "kind": "port",
        "single_thread_value": 1.0 / t_single,
        "sample": f"oracle apply chain, each stage timed on sampled output rows and scaled: "
                  f"D, E, E^T, D^T sweeps over {', '.join(nm[:4])} rows and the elementwise "
                  f"stages over {nm[4].split()[1]} rows at {nthreads} OpenMP threads "
                  f"({t_multi:.0f} s per block-apply); at 1 thread {', '.join(n1[:4])} rows "
                  f"({t_single:.0f} s per block-apply)",
    }


def cpu_baseline_cyclic(inst, sector, b, budget_s=20.0):
    """The oracle's apply chain of one cyclic eliminated sector (the paper
    recipe), whole block-applies repeated for about budget_s, at one host
    thread and at all (at most 16)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_ctypes as oc

    L = oc.lib()
    nthreads = min(16, os.cpu_count() or 1)
    p = inst["modulus"]
    o = oc.cyclic_op(inst, sector)
    dr, dc, v = oc.reference_random_state(1, sector, p, o.rows, o.cols, b)
    _, pairs, _ = oc.rowmix_plan(10001, "eliminated", sector, o.rows, p, 2)
    o.set_preconditioners(dr, dc, 2, pairs)

    def rate(threads, budget):
        L.orc_set_threads(threads)
        x, n, t0 = v, 0, time.perf_counter()
        while True:
            x = o.apply_square(x, b)
            n += 1
            el = time.perf_counter() - t0
            if el >= budget:
                return n / el, n

    r1, n1 = rate(1, budget_s * 0.4)
    rm, nm = rate(nthreads, budget_s * 0.6)
    return {"value": rm, "unit": "block-applies/s", "cores": nthreads, "kind": "port",
            "single_thread_value": r1,
            "sample": f"oracle apply_square of sector {sector} ({o.rows} x {o.cols}, b={b}): "
                      f"{nm} whole block-applies at {nthreads} OpenMP threads, {n1} at 1 thread"}


def traffic_from_profiles(workload):
    """HBM bytes per launch of the dominant kernel from the committed counter
    summary profiles/traffic_<workload>.txt (`key value` lines: FETCH_SIZE and
    WRITE_SIZE KiB per dispatch of the forward and transpose sweeps and the
    copy-kernel calibration factors), or None."""
    path = ROOT / "profiles" / f"traffic_{workload}.txt"
    if not path.exists():
        return None
    kv = {}
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) == 2 and not line.startswith("#"):
            kv[parts[0]] = float(parts[1])
    need = ("fetch_kib_fwd", "write_kib_fwd", "fetch_kib_trn", "write_kib_trn", "cal_fetch", "cal_write")
    if not all(k in kv for k in need):
        return None
    per = [(kv[f"fetch_kib_{d}"] * kv["cal_fetch"] + kv[f"write_kib_{d}"] * kv["cal_write"]) * 1024
           for d in ("fwd", "trn")]
    return sum(per) / 2


def sector_macs(spec, b):
    """GF(p) multiply-adds of one block-apply (four sweeps of equal work)."""
    return 4 * macs_per_sweep(spec.n, spec.k, spec.m_src, spec.m_tgt, b)


def bench_batch(args):
    """64 uneven sectors (seed 3): `--rounds` sequence rounds of every sector,
    all in one batch (grouped launches, graph-replayed rounds) against the same
    batch without grouping, with its grouped sweeps as persistent launches,
    the sectors run one at a time, and the batch with tile configurations
    autotuned per sector shape.  Results of the five runs must be identical."""
    import prym_green_ancillary_amd as pg
    from prym_green_ancillary_amd import workloads as wl

    jobs = wl.uneven_sectors(seed=3, count=args.sectors)
    ctx = pg.Context(0)
    dev_name, _ = ctx.device_info()
    t_setup = time.perf_counter()
    ops, vs = [], []
    for j in jobs:
        dr, dc, v, pairs = wl.sector_state(j, 3)
        T, E = j.spec.tables()
        ops.append(pg.Operator.synthetic(ctx, j.spec, j.block, dr, dc, 2, pairs, tables=(T, E)))
        vs.append(v)
    t_setup = time.perf_counter() - t_setup
    max_len = 1 + 4 * args.rounds
    applies = 2 * args.rounds * len(jobs)
    macs = sum(2 * args.rounds * sector_macs(j.spec, j.block) for j in jobs)

    def run(sets, grouped, autotune=False, persistent=False):
        out, secs, add_secs = {}, 0.0, 0.0
        for g in sets:
            batch = pg.Batch(ctx, use_graph=True, grouped=grouped, autotune=autotune, persistent=persistent)
            t0 = time.perf_counter()
            ids = [(i, batch.add(ops[i], vs[i], max_len)) for i in g]
            add_secs += time.perf_counter() - t0
            t0 = time.perf_counter()
            while batch.step(1 << 20):
                pass
            secs += time.perf_counter() - t0
            for i, jid in ids:
                out[i] = batch.result(jid)
            groups = batch.groups()
            batch.close()
        return out, secs, add_secs, groups

    # warm the code paths once (untimed)
    run([[0]], True)
    everything = [list(range(len(jobs)))]
    together, t_batch, _, _ = run(everything, True)
    ungrouped, t_ungrouped, _, _ = run(everything, False)
    persistent, t_persistent, _, _ = run(everything, True, persistent=True)
    alone, t_alone, _, _ = run([[i] for i in range(len(jobs))], False)
    # last: the autotuner keeps its choices in the operators
    tuned, t_tuned, t_tune_adds, tuned_groups = run(everything, True, autotune=True)
    for i in range(len(jobs)):
        for other in (together, ungrouped, persistent, tuned):
            if not ((other[i][0] == alone[i][0]).all() and (other[i][1] == alone[i][1]).all()):
                print(f"sector {i}: batched and lone results differ: no number reported",
                      file=sys.stderr)
                sys.exit(1)
    Ns = [j.cols for j in jobs]
    line = {
        "metric": "GF(p) block-applies/s over a batch of uneven sectors, bit-exact vs lone runs",
        "value": round(applies / t_batch, 3),
        "unit": "block-applies/s",
        "n_gpus": 1, "steps": args.rounds, "warmup": 1,
        "ms_per_step": round(1000.0 * t_batch / args.rounds, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8-limb/int32 (u32 residues mod 65521)", "data": "synthetic",
        "config": {"workload": f"{len(jobs)} uneven sectors, seed 3: n 12..22, k 3..6, M 20..300, "
                               f"M'=M+ceil(M/10), p=65521, b in {{8,16}}, N {min(Ns)}..{max(Ns)}",
                   "sectors": len(jobs), "rounds_per_sector": args.rounds, "sum_N": sum(Ns),
                   "parallelism": "single GPU, one stream, one launch per stage per group of "
                                  "compatible sectors, one HIP graph per round"},
        "gf_macs_per_s": float(f"{macs / t_batch:.4g}"),
        "one_at_a_time_block_applies_per_s": round(applies / t_alone, 3),
        "ungrouped_block_applies_per_s": round(applies / t_ungrouped, 3),
        "persistent_sweeps_block_applies_per_s": round(applies / t_persistent, 3),
        "autotuned": {"block_applies_per_s": round(applies / t_tuned, 3),
                      "add_s_with_tuning": round(t_tune_adds, 2),
                      "groups_and_grouped_jobs_at_end": list(tuned_groups)},
        "batch_vs_one_at_a_time": round(t_alone / t_batch, 3),
        "bit_exact_batched_vs_lone": True,
        "setup_s": round(t_setup, 2),
        "device": dev_name,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_serve(args):
    """Poisson stream (seed 5, 2 jobs/s) replayed in real time into the
    serving loop; each job runs its sequence terms (to the full length 2N/b
    + 8 when that is at most --full-max, so that its rank stage can finish;
    the others stop at --max-len) and, unless --no-rank, the host rank stage
    on its sequence.  Reports
    submit-to-finish latency p50/p99 and throughput, against the same stream
    through a server that runs one job at a time (and, with --plain-baseline,
    one job at a time on the plain u64 path)."""
    import prym_green_ancillary_amd as pg
    from prym_green_ancillary_amd import workloads as wl

    jobs = wl.poisson_stream(seed=5, count=args.jobs, rate=2.0)
    ctx = pg.Context(0)
    dev_name, total = ctx.device_info()
    prepared = []
    t_setup = time.perf_counter()
    for j in jobs:
        dr, dc, v, pairs = wl.sector_state(j, 5)
        T, E = j.spec.tables()
        desc = pg.Operator.synthetic_desc(j.spec, j.block, pg.METHOD_AUTO, T, E)
        full = j.full_length()
        L = full if full <= args.full_max else min(full, args.max_len)
        prepared.append((j, desc, (T, E), dr, dc, v, pairs, L))
    plain_descs = [pg.Operator.synthetic_desc(j.spec, j.block, pg.METHOD_PLAIN, *keep)
                   for (j, _, keep, *_rest) in prepared] if args.plain_baseline else None
    t_setup = time.perf_counter() - t_setup
    budget = int(args.hbm_fraction * total)

    def replay(max_active, descs=None):
        srv = pg.Server(ctx, budget, max_active=max_active, replays_per_step=1)
        if not args.no_rank:
            srv.enable_rank(threads=args.rank_threads)
        t0 = time.perf_counter()
        tickets = []
        for q, (j, desc, keep, dr, dc, v, pairs, L) in enumerate(prepared):
            if descs is not None:
                desc = descs[q]
            delay = j.arrival - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(desc, dr, dc, v, L, 2, pairs, keep=keep))
        res = []
        for t in tickets:
            if srv.wait(t, 1200.0) != pg.Server.DONE:
                print("a serving job failed: no number reported", file=sys.stderr)
                sys.exit(1)
            rk = srv.rank(t)[::2] if not args.no_rank else (None, 0.0)
            res.append(srv.result(t) + rk)
        wall = time.perf_counter() - t0
        stats = srv.stats()
        srv.close()
        lat = [r[2][4] - r[2][0] for r in res]
        return res, wall, lat, stats

    def phases(res):
        t = np.array([r[2] for r in res])
        d = np.diff(t, axis=1).mean(axis=0)
        return {"queued": round(float(d[0]), 3), "setup": round(float(d[1]), 3),
                "in_batch": round(float(d[2]), 3), "results": round(float(d[3]), 3)}

    mixed, wall, lat, stats = replay(args.max_active)
    alone, wall1, lat1, _ = replay(1)
    others = [alone]
    if plain_descs is not None:
        plain, wallp, latp, _ = replay(1, plain_descs)
        others.append(plain)
    for other in others:
        for i, (a, b1) in enumerate(zip(mixed, other)):
            if not ((a[0] == b1[0]).all() and (a[1] == b1[1]).all()) or a[3] != b1[3]:
                print(f"job {i}: results differ between the mix and a one-at-a-time run", file=sys.stderr)
                sys.exit(1)
    applies = sum(2 * ((L - 1 + 3) // 4) for (_, _, _, _, _, _, _, L) in prepared)
    Ns = [j.cols for j in jobs]
    is_full = [L >= j.full_length() for (j, _, _, _, _, _, _, L) in prepared]

    def rank_summary(sel):
        st = [r[3].status for r, f in zip(mixed, is_full) if f == sel]
        ranks = [int(r[3].rank_lower_bound) for r, f in zip(mixed, is_full) if f == sel]
        return {"jobs": len(st),
                "status_counts": {str(k): int(v) for k, v in zip(*np.unique(st, return_counts=True))},
                "ranks": ranks if sel else None}
    line = {
        "metric": "sector job latency p50/p99 + throughput under a Poisson stream; results "
                  "identical alone vs in the mix",
        "value": round(float(np.percentile(lat, 99)), 3),
        "unit": "s (p99 submit-to-finish)",
        "n_gpus": 1, "steps": len(jobs), "warmup": 0,
        "ms_per_step": round(1000.0 * wall / len(jobs), 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8-limb/int32 (u32 residues mod 65521 or 2^31-1)", "data": "synthetic",
        "config": {"workload": f"Poisson stream seed 5, 2 jobs/s, {len(jobs)} jobs, N log-uniform "
                               f"1e4..3e7 ({min(Ns)}..{max(Ns)}), p in {{65521, 2^31-1}}, "
                               f"b in {{8,16}}, terms 2N/b+8 when that is at most {args.full_max}, "
                               f"else {args.max_len}",
                   "full_length_jobs": int(sum(is_full)),
                   "jobs": len(jobs), "hbm_budget_bytes": budget, "max_active": args.max_active,
                   "parallelism": "single GPU, continuous batching"},
        "latency_s": {"p50": round(float(np.percentile(lat, 50)), 3),
                      "p99": round(float(np.percentile(lat, 99)), 3)},
        "jobs_per_s": round(len(jobs) / wall, 3),
        "block_applies_per_s": round(applies / wall, 3),
        "mean_phase_s": phases(mixed),
        "one_at_a_time": {"mean_phase_s": phases(alone),
                          "latency_p50_s": round(float(np.percentile(lat1, 50)), 3),
                          "latency_p99_s": round(float(np.percentile(lat1, 99)), 3),
                          "jobs_per_s": round(len(jobs) / wall1, 3)},
        "plain_one_at_a_time": None if plain_descs is None else {
            "mean_phase_s": phases(plain),
            "latency_p50_s": round(float(np.percentile(latp, 50)), 3),
            "latency_p99_s": round(float(np.percentile(latp, 99)), 3),
            "jobs_per_s": round(len(jobs) / wallp, 3)},
        "peak_active": stats["peak_active"], "peak_bytes": stats["peak_bytes"],
        "bit_exact_mix_vs_alone": True,
        "rank_stage": None if args.no_rank else {
            "threads_per_job": args.rank_threads,
            "mean_s": round(float(np.mean([r[4] for r in mixed])), 4),
            "max_s": round(float(np.max([r[4] for r in mixed])), 4),
            "status_legend": {"0": "running (needs more terms)", "1": "success: rank final",
                              "2": "exceeds dimension", "3": "sequence limit"},
            "full_length": rank_summary(True),
            "capped": rank_summary(False),
        },
        "setup_s": round(t_setup, 2),
        "device": dev_name,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_paper(args):
    """The reference's own instances (tests/golden/g<genus>.json) with its
    paper recipe: b = 4, Dr/Dc/V from seed 1, two row-mixing rounds from seed
    10001, 2N/4 + 50 terms.  Every sector's sequence loop (pgw_sequence_run:
    V upload, the captured round replayed, sequence and final block back to
    the host) is timed on the wall clock; the median over the sectors is the
    quantity the reference publishes for these instances (median base-sector
    sequence-loop time, RESULTS.md:26-31; other hardware, same recipe).

    --paper-terms T runs only the first T terms of each loop (a sector whose
    whole loop outlasts a run) and reports the loop time projected to the full
    length from the measured rounds, marked as such (`projected`).

    --eager-rounds R instead runs R eager rounds (apply, Gram projection)
    on the first sector, without graphs, for kernel traces."""
    import prym_green_ancillary_amd as pg

    gold = ROOT / "tests" / "golden"
    inst = json.loads((gold / f"g{args.genus}.json").read_text())
    case = next(c for c in json.loads((gold / "cases.json").read_text())["cases"]
                if c["genus"] == args.genus)
    p, b = inst["modulus"], 4
    sectors = (list(range(case["order"])) if not args.paper_sectors
               else [int(x) for x in args.paper_sectors.split(",")])
    ctx = pg.Context(0)
    dev_name, _ = ctx.device_info()
    loops, per_sector = [], []
    roofline = None
    for s in sectors:
        d, keep = pg.Operator.cyclic_desc(inst, s, b)
        rows, cols = pg.Operator.dims(d)
        dr, dc, v = pg.reference_random_state(1, s, p, rows, cols, b)
        pairs = pg.rowmix_plan(pg.rowmix_realized_seed(10001, "eliminated", s, rows, p), rows, 2, p)
        op = pg.Operator.cyclic(ctx, inst, s, b, dr, dc, 2, pairs)
        if args.eager_rounds:
            import torch
            x = torch.from_numpy(v.view(np.int32).copy()).cuda()
            y = torch.empty_like(x)
            g = torch.zeros(b * (b + 1) // 2, dtype=torch.int32, device="cuda")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.eager_rounds):
                op.apply_square(x.data_ptr(), y.data_ptr())
                op.gram_upper(x.data_ptr(), y.data_ptr(), g.data_ptr())
                x, y = y, x
            ctx.synchronize()
            us = (time.perf_counter() - t0) / args.eager_rounds * 1e6
            print(json.dumps({"eager": f"g{args.genus} sector {s}", "rows": rows, "cols": cols,
                              "rounds": args.eager_rounds, "us_per_apply_and_gram": round(us, 2)}),
                  flush=True)
            op.close()
            return
        max_len = 2 * cols // b + 50
        full_rounds = (pg.sequence_length(max_len) - 1) // 4
        run_len = min(max_len, args.paper_terms) if args.paper_terms else max_len
        t0 = time.perf_counter()
        seq, _ = op.sequence(v, run_len)
        t = time.perf_counter() - t0
        ran_rounds = (seq.shape[0] - 1) // 4
        measured = t
        if ran_rounds < full_rounds:
            t = t * full_rounds / ran_rounds  # the rounds of one loop all cost the same
        if roofline is None:
            roofline = sparse_roofline(op, v, s)
        op.close()
        applies = (seq.shape[0] - 1) // 2
        loops.append(t)
        per_sector.append({"sector": s, "rows": rows, "cols": cols, "terms": int(seq.shape[0]),
                           "loop_s": round(t, 4), "block_applies_per_s": round(applies / measured, 1),
                           "measured_s": round(measured, 4), "projected": ran_rounds < full_rounds,
                           "full_terms": pg.sequence_length(max_len)})
        print(f"g{args.genus} sector {s}: {rows} x {cols}, {seq.shape[0]} terms in {measured:.3f} s"
              + (f" (whole loop projected: {t:.1f} s)" if ran_rounds < full_rounds else ""),
              file=sys.stderr, flush=True)
    med = float(np.median(loops))
    pub = case["published_base_gpu_time_median_seconds"]
    cpu = None if args.no_cpu_baseline else cpu_baseline_cyclic(inst, sectors[0], b, args.cpu_budget)
    line = {
        "metric": f"median base-sector sequence-loop time, genus {args.genus} (the reference's "
                  "published metric for its paper instances)",
        "value": round(med, 4), "unit": "s", "n_gpus": 1, "steps": len(sectors), "warmup": 0,
        "ms_per_step": round(1000.0 * med, 2), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": round(med / pub, 4),
        "dtype": f"u32 residues mod {p} (exact u64 sums)", "data": f"data/g{args.genus}.json of the reference",
        "config": {"workload": f"g{args.genus} eliminated base sectors {sectors[0]}..{sectors[-1]}, "
                               f"b=4, seed 1, row mixing 2 rounds seed 10001, 2N/4+50 terms",
                   "sectors": len(sectors), "parallelism": "single GPU, one sector at a time"},
        "published": {"median_loop_s": pub, "gpu": case["published_gpu_model"],
                      "speedup_vs_published": round(pub / med, 3)},
        "median_block_applies_per_s": round(float(np.median([x["block_applies_per_s"] for x in per_sector])), 1),
        "per_sector": per_sector,
        "device": dev_name,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    if any(x["projected"] for x in per_sector):
        line["projected"] = {"terms_run": args.paper_terms,
                             "note": "loop times are projected to the full length from the measured rounds"}
    print(json.dumps(line), flush=True)


def sparse_roofline(op, v, sector, iters=20):
    """HBM roofline of the cyclic apply's sparse maps (k_sell_spmm, the
    dominant kernel): pgw_bench_stages times every launch of `iters` applies
    with HIP events on the operator's stream; the bytes are each map's
    algorithmic ones (entries once, input and output blocks once)."""
    import torch
    a = torch.from_numpy(v.view(np.int32).copy()).cuda()
    bb = torch.empty_like(a)
    stages = op.bench_stages(a.data_ptr(), bb.data_ptr(), iters)
    ms, by, nlaunch = stages["sparse_maps"]
    gbs = by / (ms * 1e-3) / 1e9
    peak = 8000.0  # GB/s, MI355X HBM3E
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": peak, "unit": "GB/s",
            "frac": round(gbs / peak, 4), "traffic": None, "kernel": "k_sell_spmm",
            "sector": sector, "ms_per_apply": round(ms, 4), "launches_per_apply": nlaunch,
            "bytes_per_apply": by}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--loop-rounds", type=int, default=8,
                    help="sector workloads: rounds of the timed sequence loop (4 terms each)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--workload", choices=["sector", "large", "batch", "serve", "paper"], default="large",
                    help="large: the N=53.8M sector (default); sector: the mid-size sector; "
                         "batch: 64 uneven sectors; serve: Poisson job stream; paper: the "
                         "reference's g20/g22 instances")
    ap.add_argument("--genus", type=int, default=22, choices=[20, 22, 24, 26, 28], help="paper: instance")
    ap.add_argument("--paper-sectors", default="", help="paper: comma-separated sectors (default all)")
    ap.add_argument("--paper-terms", type=int, default=0,
                    help="paper: run only this many terms per loop and project the full loop time")
    ap.add_argument("--eager-rounds", type=int, default=0,
                    help="paper: time this many eager apply+Gram rounds on the first sector instead")
    ap.add_argument("--rounds", type=int, default=4, help="batch: sequence rounds per sector")
    ap.add_argument("--sectors", type=int, default=64, help="batch: sectors in the mix")
    ap.add_argument("--jobs", type=int, default=24, help="serve: jobs in the stream")
    ap.add_argument("--max-len", type=int, default=33,
                    help="serve: sequence-term cap of the jobs that do not run to full length")
    ap.add_argument("--full-max", type=int, default=16384,
                    help="serve: jobs whose full length 2N/b+8 is at most this run to full length")
    ap.add_argument("--max-active", type=int, default=16, help="serve: concurrent jobs")
    ap.add_argument("--hbm-fraction", type=float, default=0.5, help="serve: HBM budget share")
    ap.add_argument("--no-rank", action="store_true", help="serve: skip the host rank stage")
    ap.add_argument("--rank-threads", type=int, default=8, help="serve: rank-stage threads per job")
    ap.add_argument("--plain-baseline", action="store_true",
                    help="serve: also replay the stream one job at a time on the plain u64 path")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs one GPU only (asked for --gpus {args.gpus}); the sector path "
              "has no multi-GPU mode in this build", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "paper":
        return bench_paper(args)

    import torch
    import prym_green_ancillary_amd as pg

    w = WORKLOADS[args.workload]
    n, k, M, Mp, p, b = w["n"], w["k"], w["M"], w["M_prime"], w["prime"], w["block"]
    spec = pg.SyntheticSpec(n, k, M, Mp, p, w["table_seed"], w["table_seed"] + 1)
    T, E = spec.tables()
    N = spec.cols
    dr, dc, v = pg.reference_random_state(w["seed"], 0, p, N, N, b)
    realized = pg.rowmix_realized_seed(10001, "synthetic", 0, N, p)
    pairs = pg.rowmix_plan(realized, N, 2, p)

    ctx = pg.Context(0)
    dev_name, _ = ctx.device_info()
    op = pg.Operator.synthetic(ctx, spec, b, dr, dc, 2, pairs, method=pg.METHOD_MFMA, tables=(T, E))
    # the one-pass row mixing (untimed host work, once per operator; the
    # plain operator keeps the pair passes, so the bit-exactness gate below
    # also compares the two)
    t_prep = time.perf_counter()
    op.prepare_row_mixing()
    prep_s = time.perf_counter() - t_prep
    plain = pg.Operator.synthetic(ctx, spec, b, dr, dc, 2, pairs, method=pg.METHOD_PLAIN, tables=(T, E))
    x = torch.from_numpy(v.view(np.int32).copy()).cuda()
    ya, yb = torch.empty_like(x), torch.empty_like(x)
    torch.cuda.synchronize()
    # the autotuner picks the MFMA tile configuration of each sweep (untimed)
    tile_choice, tile_ms = op.autotune(3)

    # bit-exactness gate: one apply on each path from the same input
    op.apply_square(x.data_ptr(), ya.data_ptr(), None)
    plain.apply_square(x.data_ptr(), yb.data_ptr(), None)
    ctx.synchronize()
    if not torch.equal(ya, yb):
        print("MFMA and plain block-applies differ: no number reported", file=sys.stderr)
        sys.exit(1)

    # warmup, then K timed block-applies (ping-pong), HIP events on the op's stream
    a, bb = x.clone(), torch.empty_like(x)
    if args.warmup:
        op.bench_apply(a.data_ptr(), bb.data_ptr(), args.warmup)
    ctx.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    total_ms, sweep_ms, sweeps = op.bench_apply(a.data_ptr(), bb.data_ptr(), args.steps)
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = total_ms / args.steps
    value = 1000.0 / ms_per_step

    # per-apply latency distribution (each apply synchronised on its own)
    lat = []
    for i in range(min(args.steps, 20)):
        t1 = time.perf_counter()
        op.apply_square(a.data_ptr(), bb.data_ptr(), None)
        ctx.synchronize()
        lat.append((time.perf_counter() - t1) * 1000.0)
        a, bb = bb, a

    # the sequence loop itself (pgw_sequence_run: V upload, V^T V, captured
    # rounds of two applies and four Gram projections, results back), the
    # quantity the reference times; block-applies/s with the projections in
    loop_len = 1 + 4 * max(1, args.loop_rounds)
    t0 = time.perf_counter()
    op.sequence(v, loop_len)
    loop_s = time.perf_counter() - t0
    loop_rate = (loop_len - 1) / 2 / loop_s

    # plain u64 path on the same workload
    plain_ms, _, _ = plain.bench_apply(a.data_ptr(), bb.data_ptr(), max(1, args.plain_steps))
    plain_rate = 1000.0 / (plain_ms / max(1, args.plain_steps))

    nl = 4 if p > 65536 else 2
    ops_per_sweep = 2.0 * macs_per_sweep(n, k, M, Mp, b) * nl * nl  # int8 ops, algorithmic
    achieved = ops_per_sweep / (sweep_ms * 1e-3) / 1e12
    peak = 5000.0  # TOP/s dense int8, MI355X
    traffic = traffic_from_profiles(args.workload)
    roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TOP/s",
                "frac": round(achieved / peak, 4),
                "traffic": None if traffic is None else float(f"{traffic:.4g}"),
                "kernel": "k_dense_mfma", "sweep_ms": round(sweep_ms, 4),
                "sweeps_per_apply": sweeps, "tile_configs": tile_choice,
                "tile_ms": [round(t, 4) for t in tile_ms],
                "ops_per_launch": ops_per_sweep}

    # the HBM-bound stages of the apply, HIP events around every launch
    stages = op.bench_stages(a.data_ptr(), bb.data_ptr(), 3)
    hbm_peak = 8000.0  # GB/s, MI355X HBM3E
    hbm = {}
    for name in ("digit_split", "row_mixing"):
        if name in stages:
            ms, by, nlaunch = stages[name]
            gbs = by / (ms * 1e-3) / 1e9
            hbm[name] = {"ms_per_apply": round(ms, 4), "launches": nlaunch, "bytes": by,
                         "GB/s": round(gbs, 1), "frac": round(gbs / hbm_peak, 4)}
    if hbm:
        ms_all = sum(h["ms_per_apply"] for h in hbm.values())
        by_all = sum(h["bytes"] for h in hbm.values())
        hbm["elementwise_total"] = {"ms_per_apply": round(ms_all, 4), "bytes": by_all,
                                    "GB/s": round(by_all / (ms_all * 1e-3) / 1e9, 1),
                                    "frac": round(by_all / (ms_all * 1e-3) / 1e9 / hbm_peak, 4),
                                    "share_of_apply": round(ms_all / ms_per_step, 4)}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(spec, T, E, dr, dc, pairs, v, b, budget_s=args.cpu_budget)

    line = {
        "metric": "GF(p) block-applies/s + sector p99 latency; int8-limb MFMA vs plain, bit-exact",
        "value": round(value, 3),
        "unit": "block-applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8-limb/int32 (u32 residues mod 2^31-1)",
        "data": "synthetic",
        "config": {"workload": f"{w['name']}: n={n} k={k} M={M} M'={Mp} N={N} "
                               f"p={'2^31-1' if p == P31 else p} b={b} rowmix 2 rounds",
                   "n": n, "k": k, "M": M,
                   "M_prime": Mp, "N": N, "prime": p, "block": b, "parallelism": "single GPU"},
        "sequence_loop": {"terms": loop_len, "s": round(loop_s, 3),
                          "block_applies_per_s": round(loop_rate, 3),
                          "note": "pgw_sequence_run wall time: V upload and final copy-back "
                                  "included, Gram projections included"},
        "plain_block_applies_per_s": round(plain_rate, 3),
        "mfma_vs_plain": round(value / plain_rate, 3),
        "bit_exact_mfma_vs_plain": True,
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                       "p99": round(float(np.percentile(lat, 99)), 4)},
        "wall_s_timed": round(wall, 3),
        "device": dev_name,
        "roofline": roofline,
        "hbm_stages": hbm,
        "row_mixing": {"form": "one pass along the pairing chains (MFMA operator); pair passes (plain operator)",
                       "prepare_s": round(prep_s, 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Block-apply throughput of one sector operator on one MI355X.

Workload: the single mid-size sector (n=20, k=5, M=200, M'=220,
N = C(20,5)*200 = 3,100,800 rows, p = 2^31-1, 16-wide uint32 block, V from
seed 2, two rounds of row mixing).  One step = one block-apply X -> S X with
S = Dc F^T P^T Dr P F Dc, F = E o D, on the int8-limb MFMA path; inputs are
resident in HBM before the timed region.  The plain u64 path is timed beside
it on the same input, and the two outputs must be bit-identical before any
number is printed.

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
WORKLOAD = dict(n=20, k=5, M=200, M_prime=220, prime=P31, block=16, table_seed=0, seed=2)


def macs_per_sweep(n, k, M, Mp, b):
    # every entry of every output panel (b columns) sums incidences x input-panel-rows
    # products: C(n,k-1) * (n-k+1) * M' * M * b, the same for all four sweeps
    return comb(n, k - 1) * (n - k + 1) * Mp * M * b


def cpu_baseline(spec, T, E, dr, dc, pairs, v, b, budget_s=20.0):
    """Oracle D-sweep over a bounded prefix of target rows, extrapolated to a
    whole block-apply (4 sweeps of equal work)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_ctypes as oc

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = oc.synthetic_op(spec.n, spec.k, spec.m_src, spec.m_tgt, spec.prime, T, E, dr, dc, 2,
                        pairs.reshape(-1))
    R = oc.lib().orc_sweep_rows(o.h, 0)
    out = np.zeros(R * b, np.uint32)
    x = np.ascontiguousarray(v)
    rows = max(64, R // 2000)
    elapsed, done = 0.0, 0
    while elapsed < budget_s and done < R:
        r1 = min(R, done + rows)
        t0 = time.perf_counter()
        oc.lib().orc_sweep_partial(o.h, 0, x, out, b, done, r1)
        elapsed += time.perf_counter() - t0
        done = r1
        rows *= 2
    sec_per_apply = 4.0 * elapsed * (R / done)
    return {
        "value": 1.0 / sec_per_apply,
        "unit": "block-applies/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle D sweep over {done} of {R} target rows ({elapsed:.1f} s, {threads} "
                  f"OpenMP threads), x{R / done:.1f} rows and x4 sweeps per block-apply",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs one GPU only (asked for --gpus {args.gpus}); the sector path "
              "has no multi-GPU mode in this build", file=sys.stderr)
        sys.exit(2)

    import torch
    import prym_green_ancillary_amd as pg

    w = WORKLOAD
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

    # plain u64 path on the same workload
    plain_ms, _, _ = plain.bench_apply(a.data_ptr(), bb.data_ptr(), max(1, args.plain_steps))
    plain_rate = 1000.0 / (plain_ms / max(1, args.plain_steps))

    nl = 4 if p > 65536 else 2
    ops_per_sweep = 2.0 * macs_per_sweep(n, k, M, Mp, b) * nl * nl  # int8 ops, algorithmic
    achieved = ops_per_sweep / (sweep_ms * 1e-3) / 1e12
    peak = 5000.0  # TOP/s dense int8, MI355X
    roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": None,
                "kernel": "k_dense_mfma", "sweep_ms": round(sweep_ms, 4),
                "sweeps_per_apply": sweeps, "tile_configs": tile_choice,
                "tile_ms": [round(t, 4) for t in tile_ms]}

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
        "config": {"workload": "single mid sector c2: n=20 k=5 M=200 M'=220 N=3100800 "
                               "p=2^31-1 b=16 rowmix 2 rounds", "n": n, "k": k, "M": M,
                   "M_prime": Mp, "N": N, "prime": p, "block": b, "parallelism": "single GPU"},
        "plain_block_applies_per_s": round(plain_rate, 3),
        "mfma_vs_plain": round(value / plain_rate, 3),
        "bit_exact_mfma_vs_plain": True,
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                       "p99": round(float(np.percentile(lat, 99)), 4)},
        "wall_s_timed": round(wall, 3),
        "device": dev_name,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

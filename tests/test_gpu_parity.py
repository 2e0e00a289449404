"""GPU parity: the HIP path against the oracle on the same seeded inputs.

FP64 exact paths (plain one-configuration form, batched EXACT mode, table
evaluation) must be bit-identical to the oracle.  The MFMA modes are checked
against the tolerance stated per test.
"""
import sys

import numpy as np
import pytest

import cdet_amd
import helpers

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_ctx_cache = {}


def gpu_ctx(name, **over):
    key = (name, tuple(sorted(over.items())))
    if key not in _ctx_cache:
        s = dict(helpers.SETS[name])
        op = helpers.params_of(name)
        p = cdet_amd.make_params(s["N1"], s["N2"], s["nc"], s["beta"], s["U"], s["mu"], s["alpha_up"],
                                 s["alpha_down"], derived=(op.scale, op.diag_up_shift, op.diag_down_shift),
                                 **over)
        ctx = cdet_amd.Context(p)
        ctx.upload_tables(helpers.tables_of(name))
        _ctx_cache[key] = ctx
    return _ctx_cache[key]


def oracle_for(name, **over):
    from oracle import oracle as O
    s = helpers.SETS[name]
    p = O.make_params(s["N1"], s["N2"], s["nc"], s["beta"], s["U"], s["mu"], s["alpha_up"], s["alpha_down"],
                      **{{"tol_fpm": "tol_fpm"}.get(k, k): v for k, v in over.items()})
    return O.Oracle(p, helpers.tables_of(name))


def bits_equal(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


@pytest.mark.parametrize("name", ["ref2x2", "ref2x2_eq", "c1"])
def test_plain_all_subsets_bit_identical(gpu_available, name):
    ctx = gpu_ctx(name)
    orc = helpers.oracle_of(name)
    orders = [2, 3, 4, 5, 6, 7, 7]
    for k, v in enumerate(helpers.seeded_configs(1000 + len(name), orders)):
        say(f"plain all-subset {name} config {k} order {len(v)}")
        got = ctx.evaluate_all_subset_integrands(v)
        want = orc.all_subsets(v)
        assert bits_equal(got, want), (name, k, np.max(np.abs(got - want)))


@pytest.mark.parametrize("name", ["ref2x2", "c1"])
def test_plain_single_bit_identical(gpu_available, name):
    ctx = gpu_ctx(name)
    orc = helpers.oracle_of(name)
    orders = [1, 1, 2, 3, 4, 5, 6, 7]
    for k, v in enumerate(helpers.seeded_configs(2000 + len(name), orders)):
        say(f"plain single {name} config {k} order {len(v)}")
        got = ctx.evaluate_integrand(v)
        want = orc.single(v)
        assert bits_equal(got, want), (name, k, got, want)


def test_reference_envelopes_on_gpu(gpu_available):
    # the reference's all-subset vs single-order test, run through the GPU path
    from oracle import oracle as O
    ctx = gpu_ctx("ref2x2")
    u = O.uniform_stream(0x4DCD173B, 0.01, 0.99, 5 * 6 * 4).reshape(5, 6, 4)
    for env in u:
        say("reference envelope")
        allv = ctx.evaluate_all_subset_integrands(env)
        for m in range(1, 64):
            sel = env[[i for i in range(6) if m >> i & 1]]
            d = ctx.evaluate_integrand(sel)
            assert abs(allv[m] - d) <= 2e-9 + 2e-9 * max(1.0, abs(allv[m]), abs(d))


def run_batch(ctx, configs, mode, single=False, graph=False, repeats=1, kind=None):
    import torch
    orders = [len(v) for v in configs]
    plan = ctx.plan(orders, mode=mode, single=single, kind=kind)
    verts = torch.tensor(np.concatenate(configs), dtype=torch.float64, device="cuda")
    out = torch.empty(plan.out_size, dtype=torch.float64, device="cuda")
    status = torch.zeros(len(orders), dtype=torch.int32, device="cuda")
    results = []
    for _ in range(repeats):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        plan.run(verts, out, status, use_graph=graph, stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        results.append(out.cpu().numpy().copy())
    plan.close()
    _, ob = cdet_amd.packed_offsets(orders, single)
    return results, ob, status.cpu().numpy()


@pytest.mark.parametrize("name", ["ref2x2", "ref2x2_eq", "c1"])
def test_batched_exact_bit_identical(gpu_available, name):
    ctx = gpu_ctx(name)
    orc = helpers.oracle_of(name)
    rng = np.random.default_rng(5)
    orders = list(rng.integers(2, 9, size=48))
    configs = helpers.seeded_configs(3000 + len(name), orders)
    say(f"batched exact {name}: {len(configs)} configs")
    (out,), ob, status = run_batch(ctx, configs, "exact")
    assert not np.any(status & 1)
    for k, v in enumerate(configs):
        want = orc.all_subsets(v)
        assert bits_equal(out[ob[k]:ob[k + 1]], want), (name, k)
    say(f"batched exact single {name}")
    orders1 = list(rng.integers(1, 9, size=32))
    configs1 = helpers.seeded_configs(4000 + len(name), orders1)
    (out1,), _, _ = run_batch(ctx, configs1, "exact", single=True)
    for k, v in enumerate(configs1):
        assert bits_equal(out1[k], orc.single(v)), (name, k)


def test_batched_graph_replays_match_eager(gpu_available):
    ctx = gpu_ctx("c1")
    orders = [2, 5, 7, 3, 8, 6, 4, 7]
    configs = helpers.seeded_configs(77, orders)
    say("graph replay c1")
    eager, ob, _ = run_batch(ctx, configs, "exact")
    replays, _, _ = run_batch(ctx, configs, "exact", graph=True, repeats=3)
    for r in replays:
        assert bits_equal(r, eager[0])


def test_forked_order_buckets_bit_identical(gpu_available):
    # A batch of many orders spreads its order buckets over several streams
    # inside the captured graph; eager runs and graph replays give the
    # oracle's values bit for bit, and so does the same batch split into
    # single-order plans (one stream each).
    import torch
    ctx = gpu_ctx("c1")
    orc = helpers.oracle_of("c1")
    orders = [2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12] * 3 + [2, 3, 4] * 20
    configs = helpers.seeded_configs(9100, orders)
    verts = torch.tensor(np.concatenate(configs), dtype=torch.float64, device="cuda")
    plan = ctx.plan([len(v) for v in configs], "exact")
    assert plan.streams > 1
    out = torch.empty(plan.out_size, dtype=torch.float64, device="cuda")
    sptr = torch.cuda.current_stream().cuda_stream
    say(f"forked buckets: {plan.streams} streams, eager")
    plan.run(verts, out, None, use_graph=False, stream=sptr)
    torch.cuda.synchronize()
    eager = out.cpu().numpy().copy()
    _, ob = cdet_amd.packed_offsets([len(v) for v in configs], False)
    for k, v in enumerate(configs):
        if k % 3 == 0 or len(v) >= 10:
            assert bits_equal(eager[ob[k]:ob[k + 1]], orc.all_subsets(v)), k
    for r in range(3):
        say(f"forked buckets: graph replay {r}")
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        plan.run(verts, out, None, use_graph=True, stream=sptr)
        torch.cuda.synchronize()
        assert bits_equal(out.cpu().numpy(), eager)
    plan.close()
    for n in (12, 3):
        sel = [k for k, v in enumerate(configs) if len(v) == n]
        single = ctx.plan([n] * len(sel), "exact")
        assert single.streams == 1
        o1 = torch.empty(single.out_size, dtype=torch.float64, device="cuda")
        v1 = torch.tensor(np.concatenate([configs[k] for k in sel]), dtype=torch.float64, device="cuda")
        single.run(v1, o1, None, use_graph=False, stream=sptr)
        torch.cuda.synchronize()
        got = o1.cpu().numpy().reshape(len(sel), -1)
        for j, k in enumerate(sel):
            assert bits_equal(got[j], eager[ob[k]:ob[k + 1]]), (n, k)
        single.close()


def test_batches_in_flight_on_queue_streams(gpu_available):
    # Three plans replayed round-robin on three streams of hardware queues of
    # their own (cdet_stream_create), so their launches overlap on the
    # device: every replay gives the values of the same batch run alone on
    # the context stream, bit for bit, in both FP64 modes.
    import torch
    ctx = gpu_ctx("c1")
    sets = [helpers.seeded_configs(9300 + k, orders)
            for k, orders in enumerate([[8] * 40, [6, 7, 8, 9] * 10, [10] * 12])]
    streams = [ctx.stream_create() for _ in sets]
    assert len(set(streams)) == len(sets) and all(streams)
    try:
        for mode in ("exact", "mfma64"):
            plans, verts, outs, alone = [], [], [], []
            for configs in sets:
                v = torch.tensor(np.concatenate(configs), dtype=torch.float64, device="cuda")
                p = ctx.plan([len(c) for c in configs], mode)
                o = torch.empty(p.out_size, dtype=torch.float64, device="cuda")
                p.run(v, o, None, use_graph=False, stream=ctx.stream)
                torch.cuda.synchronize()
                alone.append(o.cpu().numpy().copy())
                plans.append(p)
                verts.append(v)
                outs.append(o)
            for r in range(3):
                say(f"in flight on queue streams: {mode} round {r}")
                for o in outs:
                    o.fill_(float("nan"))
                torch.cuda.synchronize()
                for k, p in enumerate(plans):
                    p.run(verts[k], outs[k], None, use_graph=True, stream=streams[k])
                torch.cuda.synchronize()
                for k in range(len(plans)):
                    assert bits_equal(outs[k].cpu().numpy(), alone[k]), (mode, r, k)
            for p in plans:
                p.close()
            if mode == "exact":
                orc = helpers.oracle_of("c1")
                assert bits_equal(alone[2].reshape(len(sets[2]), -1)[0], orc.all_subsets(sets[2][0]))
    finally:
        torch.cuda.synchronize()
        for s in streams:
            ctx.stream_destroy(s)
    with pytest.raises(RuntimeError):
        ctx.stream_destroy(0)


def test_paired_classes_bit_identical(gpu_available):
    # A single-order batch with at least two configurations per CU runs both
    # classes of a configuration in one workgroup and combines them in place;
    # values and magnitude scales equal those of the same configurations in
    # small batches (one workgroup per class, separate combine pass) bit for
    # bit, and the oracle's on a sample; with the principal minors in the same
    # workgroup too (fused_minors_max_order) as well.
    import torch
    ctx = gpu_ctx("c1")
    orc = helpers.oracle_of("c1")
    count = 540  # >= 2 per CU on a 256-CU device
    configs = helpers.seeded_configs(9300, [6] * count)
    sptr = torch.cuda.current_stream().cuda_stream

    def run(cfgs, kind):
        plan = ctx.plan([len(v) for v in cfgs], "exact", kind=kind)
        verts = torch.tensor(np.concatenate(cfgs), dtype=torch.float64, device="cuda")
        out = torch.empty(plan.out_size, dtype=torch.float64, device="cuda")
        plan.run(verts, out, None, use_graph=False, stream=sptr)
        torch.cuda.synchronize()
        plan.close()
        return out.cpu().numpy()

    for kind in (0, 2):
        say(f"paired classes kind {kind}")
        big = run(configs, kind)
        small = np.concatenate([run(configs[i:i + 90], kind) for i in range(0, count, 90)])
        assert bits_equal(big, small)
        try:
            ctx.set_tuning({"fused_minors_max_order": 10})
            say(f"paired classes kind {kind}, fused minors")
            fused = run(configs, kind)
        finally:
            ctx.set_tuning({})
        assert bits_equal(fused, small)
        if kind == 0:
            per = 1 << 6
            for k in range(0, count, 53):
                assert bits_equal(big[k * per:(k + 1) * per], orc.all_subsets(configs[k])), k


def test_fallback_configurations_in_paired_buckets_bit_identical(gpu_available):
    # Paired order buckets up to order 10 with many pivoted-determinant
    # fallbacks: the unfused kernels queue the flagged matrices between the
    # principal minors and the recursion; the fused minors-and-recursion
    # kernel defers such a configuration to the fallback pass and a second
    # recursion pass.  At TOL_FPM = 1e-3 about a fifth of the order-8 and two
    # thirds of the order-10 configurations of this set fall back, so both
    # routes run in each bucket.  Graph replays must repeat the values (the
    # deferral words are reset by the runs that drain them), and every value
    # and status word equals the unpaired kernels' and the values the
    # oracle's, bit for bit.
    ctx = gpu_ctx("c1", tol_fpm=1e-3)
    orc = oracle_for("c1", tol_fpm=1e-3)
    configs = helpers.seeded_configs(9750, [8] * 260 + [10] * 260)
    runs = {}
    try:
        # paired from 256 configurations per bucket; unfused, then fused
        for name, tuning in (("paired", {"pair_min_per_cu": 1}),
                             ("fused", {"pair_min_per_cu": 1, "fused_minors_max_order": 10})):
            ctx.set_tuning(tuning)
            say(f"fallbacks in paired buckets, {name}: graph replays")
            outs, ob, status = run_batch(ctx, configs, "exact", graph=True, repeats=3)
            for o in outs[1:]:
                assert bits_equal(o, outs[0]), name
            runs[name] = (outs[0], status)
        ctx.set_tuning({"pair_min_per_cu": 1 << 20})  # unpaired kernels
        say("unpaired kernels, same batch")
        (ref,), _, ref_status = run_batch(ctx, configs, "exact")
    finally:
        ctx.set_tuning({})
    deferred = np.flatnonzero(ref_status & 2)
    say(f"{len(deferred)} of {len(configs)} configurations took the fallback")
    assert 0 < len(deferred) < len(configs)
    assert not np.any(ref_status & 1)
    for name, (out, status) in runs.items():
        assert bits_equal(out, ref), name
        assert np.array_equal(status, ref_status), name
    for k in sorted(set(deferred[::3].tolist()) | set(range(0, len(configs), 29))):
        assert bits_equal(ref[ob[k]:ob[k + 1]], orc.all_subsets(configs[k])), k


def test_mixed_order_runs_bit_identical(gpu_available):
    # The order buckets 2..7 of a batch, and those of orders 8..10 with fewer
    # configurations than the device has CUs, run as one run of mixed orders:
    # one launch each of the principal minors, the fallback determinants and
    # the recursion, every configuration at its own order.  At TOL_FPM = 0.9
    # many matrices take the pivoted-determinant fallback.  Values and magnitude
    # scales (kinds 0 and 2; unpaired and paired recursion kernels; eager and
    # over graph replays) and single-order values (kind 1, order 1 beside the
    # run) equal those of the same configurations in single-order batches bit
    # for bit, and the oracle's on a sample.
    ctx = gpu_ctx("c1", tol_fpm=0.9)
    orc = oracle_for("c1", tol_fpm=0.9)
    rng = np.random.default_rng(11)
    orders = [int(x) for x in rng.permutation([int(x) for x in rng.integers(2, 8, size=300)] + [8] * 12 + [9] * 10
                                             + [10] * 8)]
    configs = helpers.seeded_configs(9900, orders)

    def per_order(cfgs, kind, single=False):
        # the same configurations in single-order batches, in batch order
        res = [None] * len(cfgs)
        for n in sorted(set(len(v) for v in cfgs)):
            sel = [k for k, v in enumerate(cfgs) if len(v) == n]
            (o,), ob1, st = run_batch(ctx, [cfgs[k] for k in sel], "exact", single=single, kind=kind)
            for j, k in enumerate(sel):
                res[k] = (o[ob1[j]:ob1[j + 1]], st[j])
        return res

    try:
        for tname, tuning in (("unpaired", {"pair_min_per_cu": 1 << 20}), ("paired", {"pair_min_per_cu": 1})):
            ctx.set_tuning(tuning)
            for kind in (0, 2):
                say(f"mixed-order run, {tname}, kind {kind}: graph replays")
                outs, ob, status = run_batch(ctx, configs, "exact", graph=True, repeats=3, kind=kind)
                for o in outs[1:]:
                    assert bits_equal(o, outs[0]), (tname, kind)
                (eager,), _, st_eager = run_batch(ctx, configs, "exact", kind=kind)
                assert bits_equal(eager, outs[0]), (tname, kind)
                assert np.array_equal(st_eager, status)
                assert np.count_nonzero(status & 2) > 0
                assert not np.any(status & 1)
                say(f"mixed-order run, {tname}, kind {kind}: single-order batches")
                ref = per_order(configs, kind)
                for k in range(len(configs)):
                    assert bits_equal(eager[ob[k]:ob[k + 1]], ref[k][0]), (tname, kind, k)
                    assert status[k] == ref[k][1], (tname, kind, k)
                if kind == 0:
                    for k in range(0, len(configs), 7):
                        assert bits_equal(eager[ob[k]:ob[k + 1]], orc.all_subsets(configs[k])), (tname, k)
    finally:
        ctx.set_tuning({})
    orders1 = [int(x) for x in rng.integers(1, 11, size=96)]
    configs1 = helpers.seeded_configs(9950, orders1)
    say("mixed-order run, single-order values")
    (out1,), _, _ = run_batch(ctx, configs1, "exact", single=True)
    ref1 = per_order(configs1, None, single=True)
    for k, v in enumerate(configs1):
        assert bits_equal(out1[k], orc.single(v)), k
        assert bits_equal(out1[k:k + 1], ref1[k][0]), k


def test_contracted_lookups_exclude_diagonal_pairs(gpu_available):
    # The FP64 matrix-core mode evaluates the diagonal pairs (i, i) -- lower
    # branch at v = 1 exactly -- from the coefficient blocks' row sums while
    # planning, so its contraction takes 2 n fewer lookups per configuration
    # than the bit-identical form's, which takes every non-corner pair (no
    # corners at these times: tau >= 1e-13 for every vertex here).
    ctx = gpu_ctx("c1")
    orders = [2, 5, 6, 7, 9, 10] * 5
    configs = helpers.seeded_configs(9810, orders)
    assert min(float(np.min(v[:, 0])) for v in configs) * helpers.SETS["c1"]["beta"] > 1e-13
    counts = {}
    for mode in ("exact", "mfma64"):
        import torch
        plan = ctx.plan(orders, mode=mode)
        verts = torch.tensor(np.concatenate(configs), dtype=torch.float64, device="cuda")
        out = torch.empty(plan.out_size, dtype=torch.float64, device="cuda")
        plan.run(verts, out, None, use_graph=False, stream=torch.cuda.current_stream().cuda_stream)
        counts[mode] = plan.contracted()
        plan.close()
    say(f"contracted lookups {counts}")
    assert counts["exact"][0] == sum(2 * n * n for n in orders)
    assert counts["exact"][0] - counts["mfma64"][0] == sum(2 * n for n in orders)
    for lookups, tiles in counts.values():  # tiles of at most 64 lookups of one block
        assert 0 < tiles and lookups <= 64 * tiles


def rel_err(got, want):
    scale = np.maximum(np.abs(want), 1e-300)
    return np.abs(got - want) / scale


@pytest.mark.parametrize("name", ["ref2x2", "c1", "c1_neq"])
def test_batched_mfma64_within_tolerance(gpu_available, name):
    # FP64 matrix-core contraction: table values agree to ~1e-15 relative; the
    # integrand is checked at 1e-9 relative per mask, measured against the
    # largest |value| of the configuration for masks that cancel below it.
    ctx = gpu_ctx(name)
    orc = helpers.oracle_of(name)
    orders = [2, 3, 4, 5, 6, 7, 8, 8]
    configs = helpers.seeded_configs(5000 + len(name), orders)
    say(f"batched mfma64 {name}")
    (out,), ob, _ = run_batch(ctx, configs, "mfma64")
    worst = 0.0
    for k, v in enumerate(configs):
        want = orc.all_subsets(v)
        got = out[ob[k]:ob[k + 1]]
        scale = np.maximum(np.abs(want), 1e-6 * np.max(np.abs(want)))
        err = np.max(np.abs(got - want) / scale)
        worst = max(worst, err)
    say(f"mfma64 {name} worst scaled error {worst:.3e}")
    assert worst < 1e-9


def test_mfma64_plan_follows_table_changes(gpu_available):
    # The FP64 matrix-core contraction reads step-order copies of the tables;
    # tables uploaded after the plan (and its captured graph) exist must be
    # the ones a later run reads, eager and graph-replayed alike.
    import torch
    from oracle import oracle as O
    name = "c1"
    s = helpers.SETS[name]
    op = helpers.params_of(name)
    ctx = cdet_amd.Context(cdet_amd.make_params(s["N1"], s["N2"], s["nc"], s["beta"], s["U"], s["mu"],
                                                s["alpha_up"], s["alpha_down"],
                                                derived=(op.scale, op.diag_up_shift, op.diag_down_shift)))
    base = helpers.tables_of(name)
    ctx.upload_tables(base)
    orders = [3, 5, 6, 8]
    configs = helpers.seeded_configs(6100, orders)
    plan = ctx.plan(orders, mode="mfma64")
    verts = torch.tensor(np.concatenate(configs), dtype=torch.float64, device="cuda")
    out = torch.empty(plan.out_size, dtype=torch.float64, device="cuda")
    _, ob = cdet_amd.packed_offsets(orders, False)
    stream = torch.cuda.current_stream().cuda_stream

    def check(tables, graph):
        orc = O.Oracle(op, tables)
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        plan.run(verts, out, None, use_graph=graph, stream=stream)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        for k, v in enumerate(configs):
            want = orc.all_subsets(v)
            scale = np.maximum(np.abs(want), 1e-6 * np.max(np.abs(want)))
            assert np.max(np.abs(got[ob[k]:ob[k + 1]] - want) / scale) < 1e-9, (graph, k)

    say("mfma64 plan, first tables")
    check(base, True)
    made = {}  # keep the kinds that share a table sharing it (uploads alias them)
    for t in base:
        if id(t) not in made:
            made[id(t)] = O.Table(t.upper * (1.0 + 1e-3), t.lower * (1.0 - 1e-3), t.ucorner, t.lcorner)
    bent = [made[id(t)] for t in base]
    # a kind another kind aliases cannot be uploaded: give the aliasing kinds
    # their own copy first, upload the sources, then alias again
    src = {k: next(i for i in range(4) if base[i] is base[k]) for k in range(4)}
    for k in range(4):
        if src[k] != k:
            ctx.upload_table(k, bent[k].upper, bent[k].lower, bent[k].ucorner, bent[k].lcorner)
    for k in range(4):
        if src[k] == k:
            ctx.upload_table(k, bent[k].upper, bent[k].lower, bent[k].ucorner, bent[k].lcorner)
    for k in range(4):
        if src[k] != k:
            ctx.alias_table(k, src[k])
    say("mfma64 plan, tables uploaded after the plan")
    check(bent, True)
    check(bent, False)
    plan.close()
    ctx.close()


@pytest.mark.parametrize("name", ["c1", "c1_neq"])
def test_conditioning_scale_bounds_values_and_mfma_error(gpu_available, name):
    # The conditioning scale of a value is the sum of |a[mask]| and |terms| of
    # its subset chain, combined as |twisted| + 2 |ordinary|; it bounds |value|
    # and, times a few ulps of the operands, the error of any FP64 evaluation
    # order of the same terms.
    ctx = gpu_ctx(name)
    orders = [2, 4, 6, 8, 9]
    configs = helpers.seeded_configs(5100 + len(name), orders)
    say(f"conditioning scales {name}")
    (cond,), ob, _ = run_batch(ctx, configs, "exact", kind=2)
    (ex,), _, _ = run_batch(ctx, configs, "exact")
    (mf,), _, _ = run_batch(ctx, configs, "mfma64")
    assert np.all(np.isfinite(cond)) and np.all(cond >= 0)
    assert np.all(cond >= np.abs(ex) * (1 - 1e-12))
    worst = float(np.max(np.abs(mf - ex) / np.maximum(cond, 1e-300)))
    say(f"{name} max |mfma64 - exact| / conditioning = {worst:.3e}")
    assert worst < 1e-11


def test_split32_runs_and_reports_error(gpu_available):
    ctx = gpu_ctx("c1")
    orc = helpers.oracle_of("c1")
    configs = helpers.seeded_configs(6000, [4, 6, 8])
    say("batched split32 c1")
    (out,), ob, _ = run_batch(ctx, configs, "split32")
    assert np.all(np.isfinite(out))
    errs = []
    for k, v in enumerate(configs):
        want = orc.all_subsets(v)
        got = out[ob[k]:ob[k + 1]]
        scale = np.maximum(np.abs(want), 1e-6 * np.max(np.abs(want)))
        errs.append(np.max(np.abs(got - want) / scale))
    say(f"split32 c1 scaled errors {errs}")
    assert max(errs) < 1e-3  # split-FP32 accumulates at 2^-24 per K step


def test_fpm_fallback_path_bit_identical(gpu_available):
    # A large TOL_FPM forces the pivoted-determinant fallback
    # (include/cdet/fast_principle_minors.h:180-186) on most matrices.
    ctx = gpu_ctx("c1", tol_fpm=0.9)
    orc = oracle_for("c1", tol_fpm=0.9)
    ctx.reset_counters()
    configs = helpers.seeded_configs(7000, [3, 5, 6])
    for v in configs:
        say(f"fallback order {len(v)}")
        got = ctx.evaluate_all_subset_integrands(v)
        assert bits_equal(got, orc.all_subsets(v))
    assert ctx.counters()["fpm_fallbacks"] > 0
    assert orc.fpm_fallbacks > 0


def test_batched_fallback_repair_bit_identical(gpu_available):
    # Batched form of the fallback: matrices whose Schur pivot trips TOL_FPM
    # are queued and redone with pivoted determinants, in LDS up to order 11
    # and in global scratch above.
    ctx = gpu_ctx("c1", tol_fpm=0.9)
    orc = oracle_for("c1", tol_fpm=0.9)
    configs = helpers.seeded_configs(7100, [3, 5, 6, 8, 11, 12, 9, 11])
    say("batched fallback repair")
    (out,), ob, status = run_batch(ctx, configs, "exact")
    assert np.count_nonzero(status & 2) > 0
    for k, v in enumerate(configs):
        assert bits_equal(out[ob[k]:ob[k + 1]], orc.all_subsets(v)), k


def test_high_orders_global_scratch_bit_identical(gpu_available):
    ctx = gpu_ctx("c1")
    orc = helpers.oracle_of("c1")
    for n in (12, 13, 14):
        v = helpers.seeded_configs(8000 + n, [n])[0]
        say(f"order {n} plain")
        got = ctx.evaluate_all_subset_integrands(v)
        assert bits_equal(got, orc.all_subsets(v))
    say("orders 12-14 batched")
    configs = helpers.seeded_configs(8100, [12, 13, 14, 13])
    (out,), ob, _ = run_batch(ctx, configs, "exact")
    for k, v in enumerate(configs):
        assert bits_equal(out[ob[k]:ob[k + 1]], orc.all_subsets(v))


def test_nonfinite_tables_raise_reference_message(gpu_available):
    from oracle import oracle as O
    s = helpers.SETS["ref2x2"]
    p = cdet_amd.make_params(s["N1"], s["N2"], s["nc"], s["beta"], s["U"], s["mu"], s["alpha_up"], s["alpha_down"])
    ctx = cdet_amd.Context(p)
    good = helpers.tables_of("ref2x2")
    bad = O.Table(np.full_like(good[0].upper, np.inf), np.full_like(good[0].lower, np.inf),
                  np.full_like(good[0].ucorner, np.inf), np.full_like(good[0].lcorner, np.inf))
    ctx.upload_tables([bad, bad, good[2], good[3]])
    say("nonfinite")
    with pytest.raises(RuntimeError, match="not finite"):
        ctx.evaluate_all_subset_integrands(helpers.seeded_configs(9, [4])[0])
    with pytest.raises(RuntimeError, match="envelope order >= 2"):
        ctx.evaluate_all_subset_integrands(np.full((1, 4), 0.5))
    ctx.close()


@pytest.mark.parametrize("mode,name", [("exact", "c1"), ("mfma64", "c1"), ("mfma64", "c1_neq")])
def test_green_pair_eval(gpu_available, mode, name):
    import torch
    ctx = gpu_ctx(name)
    orc = helpers.oracle_of(name)
    rng = np.random.default_rng(3)
    n = 4096
    qi = np.zeros((n, 7), dtype=np.int32)
    qi[:, 0] = rng.integers(0, 2, n)
    qi[:, 1:3] = rng.integers(0, 2, (n, 2))
    qi[qi[:, 0] == 1, 1:3] = 0
    qi[:, 3:7] = rng.integers(0, 3, (n, 4))
    qt = rng.uniform(0, 4.0, (n, 2))
    qt[:5, 0] = 0.0
    out = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
    d_qi = torch.tensor(qi, device="cuda")
    d_qt = torch.tensor(qt, device="cuda")
    say(f"green pair eval {mode}")
    import ctypes
    rc = cdet_amd.lib().cdet_green_pair_eval(ctx.handle, ctypes.c_void_p(d_qi.data_ptr()),
                                             ctypes.c_void_p(d_qt.data_ptr()), n, ctypes.c_void_p(out.data_ptr()),
                                             cdet_amd.MODES[mode], None)
    assert rc == 0, cdet_amd._err()
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(n, 2)
    want = np.array([orc.green_pair(2 * qi[k, 0], 2 * qi[k, 0] + 1, *qi[k, 1:7], qt[k, 0], qt[k, 1])
                     for k in range(n)])
    if mode == "exact":
        assert bits_equal(got, want)
    else:
        assert np.max(np.abs(got - want)) < 1e-13


def test_green_pair_eval_large_batch_tiles(gpu_available):
    # 800,000 lookups: enough that the FP64 matrix-core contraction takes tiles
    # of up to 64 lookups (1-4 N tiles per wave) instead of 16; checked against
    # the bit-identical FP64 path (itself pinned to the oracle above) at the
    # table-level bound of the MFMA form.
    import ctypes
    import torch
    ctx = gpu_ctx("c1")
    rng = np.random.default_rng(11)
    n = 800_000
    qi = np.zeros((n, 7), dtype=np.int32)
    qi[:, 0] = rng.integers(0, 2, n)
    qi[:, 1:3] = rng.integers(0, 2, (n, 2))
    qi[qi[:, 0] == 1, 1:3] = 0
    qi[:, 3:7] = rng.integers(0, 3, (n, 4))
    qt = rng.uniform(0, 4.0, (n, 2))
    d_qi = torch.tensor(qi, device="cuda")
    d_qt = torch.tensor(qt, device="cuda")
    res = {}
    for mode in ("exact", "mfma64"):
        say(f"green pair eval, {n} lookups, {mode}")
        out = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
        rc = cdet_amd.lib().cdet_green_pair_eval(ctx.handle, ctypes.c_void_p(d_qi.data_ptr()),
                                                 ctypes.c_void_p(d_qt.data_ptr()), n,
                                                 ctypes.c_void_p(out.data_ptr()), cdet_amd.MODES[mode], None)
        assert rc == 0, cdet_amd._err()
        torch.cuda.synchronize()
        res[mode] = out.cpu().numpy()
    assert np.max(np.abs(res["mfma64"] - res["exact"])) < 1e-13
    # the stage profile of the same call (cdet_green_pair_profile) gives the
    # same values bit for bit and positive stage times
    for mode in ("exact", "mfma64"):
        out = torch.full((2 * n,), float("nan"), dtype=torch.float64, device="cuda")
        ms = (ctypes.c_float * 3)()
        rc = cdet_amd.lib().cdet_green_pair_profile(ctx.handle, ctypes.c_void_p(d_qi.data_ptr()),
                                                    ctypes.c_void_p(d_qt.data_ptr()), n,
                                                    ctypes.c_void_p(out.data_ptr()), cdet_amd.MODES[mode], None, ms)
        assert rc == 0, cdet_amd._err()
        say(f"green pair profile {mode}: stages {list(ms)} ms")
        assert np.array_equal(out.cpu().numpy().view(np.int64), res[mode].view(np.int64))
        assert all(x > 0.0 for x in ms)


def test_plain_entry_points_from_concurrent_threads(gpu_available):
    # The reference calls the integrand from one std::async thread per chain
    # (include/cdet/mcmcmc_sampler.h:1104-1110).  Four host threads (ctypes
    # releases the GIL inside the call) evaluate on one context at once; every
    # value is the oracle's, bit for bit.
    import threading
    name = "c1"
    s = helpers.SETS[name]
    op = helpers.params_of(name)
    ctx = cdet_amd.Context(cdet_amd.make_params(s["N1"], s["N2"], s["nc"], s["beta"], s["U"], s["mu"],
                                                s["alpha_up"], s["alpha_down"],
                                                derived=(op.scale, op.diag_up_shift, op.diag_down_shift)))
    ctx.upload_tables(helpers.tables_of(name))
    orc = helpers.oracle_of(name)
    per_thread = [helpers.seeded_configs(9700 + t, [7, 5, 8, 6, 7, 4, 8, 3]) for t in range(4)]
    want = [[(orc.all_subsets(v), orc.single(v)) for v in cfgs] for cfgs in per_thread]
    got = [[None] * len(c) for c in per_thread]
    errors = []
    start = threading.Barrier(4)

    def work(t):
        try:
            start.wait()
            for rep in range(3):
                for k, v in enumerate(per_thread[t]):
                    a = ctx.evaluate_all_subset_integrands(v)
                    b = ctx.evaluate_integrand(v)
                    if rep == 2:
                        got[t][k] = (a, b)
        except Exception as e:  # reported on the main thread
            errors.append(e)

    say("plain entry points from 4 threads")
    threads = [threading.Thread(target=work, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for t in range(4):
        for k in range(len(per_thread[t])):
            assert bits_equal(got[t][k][0], want[t][k][0]), (t, k)
            assert bits_equal(got[t][k][1], want[t][k][1]), (t, k)
    assert ctx.counters()["integrand_calls"] == 4 * 3 * 2 * len(per_thread[0])
    ctx.close()


def test_green_pair_scratch_follows_the_context(gpu_available):
    # The lookup scratch belongs to its context: lookups on a 3x3 context,
    # that context closed, then a few lookups on a new 8x8 context (four
    # times the key count) give the oracle's values bit for bit.
    import ctypes
    import torch
    from oracle import oracle as O
    c1 = helpers.SETS["c1"]
    op1 = helpers.params_of("c1")
    ctx = cdet_amd.Context(cdet_amd.make_params(c1["N1"], c1["N2"], c1["nc"], c1["beta"], c1["U"], c1["mu"],
                                                c1["alpha_up"], c1["alpha_down"],
                                                derived=(op1.scale, op1.diag_up_shift, op1.diag_down_shift)))
    ctx.upload_tables(helpers.tables_of("c1"))

    def lookups(cx, qi, qt):
        out = torch.zeros(2 * len(qi), dtype=torch.float64, device="cuda")
        d_qi = torch.tensor(qi, device="cuda")
        d_qt = torch.tensor(qt, device="cuda")
        rc = cdet_amd.lib().cdet_green_pair_eval(cx.handle, ctypes.c_void_p(d_qi.data_ptr()),
                                                 ctypes.c_void_p(d_qt.data_ptr()), len(qi),
                                                 ctypes.c_void_p(out.data_ptr()), cdet_amd.MODE_EXACT, None)
        assert rc == 0, cdet_amd._err()
        torch.cuda.synchronize()
        return out.cpu().numpy().reshape(-1, 2)

    r = np.random.default_rng(8)
    qi = np.zeros((4096, 7), dtype=np.int32)
    qi[:, 1:3] = r.integers(0, 2, (4096, 2))
    qi[:, 3:7] = r.integers(0, 3, (4096, 4))
    qt = r.uniform(0, c1["beta"], (4096, 2))
    say("lookups on a 3x3 context")
    lookups(ctx, qi, qt)
    ctx.close()
    s = helpers.SETS["p8x8"]
    p = cdet_amd.make_params(s["N1"], s["N2"], s["nc"], s["beta"], s["U"], s["mu"], s["alpha_up"], s["alpha_down"])
    big = cdet_amd.Context(p)
    say("8x8 context: tables on the GPU, then 64 lookups")
    big.build_tables(pbc=True, mask=s["mask"], mask_param=s["mask_param"])
    tabs = [O.Table(*big.download_table(k)) for k in range(4)]
    orc = O.Oracle(helpers.params_of("p8x8"), tabs)
    qi2 = np.zeros((64, 7), dtype=np.int32)
    qi2[:, 0] = r.integers(0, 2, 64)
    qi2[:, 1:3] = r.integers(0, 2, (64, 2))
    qi2[qi2[:, 0] == 1, 1:3] = 0
    qi2[:, 3:7] = r.integers(0, 8, (64, 4))
    qt2 = r.uniform(0, s["beta"], (64, 2))
    got = lookups(big, qi2, qt2)
    want = np.array([orc.green_pair(2 * qi2[k, 0], 2 * qi2[k, 0] + 1, *qi2[k, 1:7], qt2[k, 0], qt2[k, 1])
                     for k in range(64)])
    assert bits_equal(got, want)
    big.close()


def test_context_close_closes_its_plans(gpu_available):
    # A plan must not outlive its context: Context.close() destroys the
    # context's live plans first, and closing them again is harmless.
    import torch
    name = "c1"
    s = helpers.SETS[name]
    ctx = cdet_amd.Context(cdet_amd.make_params(s["N1"], s["N2"], s["nc"], s["beta"], s["U"], s["mu"],
                                                s["alpha_up"], s["alpha_down"]))
    ctx.upload_tables(helpers.tables_of(name))
    configs = helpers.seeded_configs(9900, [5, 6, 7])
    plan = ctx.plan([len(v) for v in configs], "exact")
    verts = torch.tensor(np.concatenate(configs), dtype=torch.float64, device="cuda")
    out = torch.empty(plan.out_size, dtype=torch.float64, device="cuda")
    say("plan run, then context close with the plan live")
    plan.run(verts, out, None, use_graph=True, stream=torch.cuda.current_stream().cuda_stream)
    ctx.close()
    assert plan._h is None
    plan.close()
    torch.cuda.synchronize()


def perturbed_tables(tables, rel, seed, coherent=False):
    """Every coefficient c moved by at most rel |c| (random sign per
    coefficient, or sign(c) for all when coherent), corners kept: each table
    value moves by at most rel sum|c_ab|, inside the error-bound model."""
    from oracle import oracle as O
    r = np.random.default_rng(seed)
    made = {}
    out = []
    for t in tables:
        if id(t) not in made:
            arrs = []
            for a in (t.upper, t.lower):
                d = np.sign(a) if coherent else r.uniform(-1.0, 1.0, a.shape)
                arrs.append(a + rel * d * np.abs(a))
            made[id(t)] = O.Table(arrs[0], arrs[1], t.ucorner, t.lcorner)
        out.append(made[id(t)])
    return out


@pytest.mark.parametrize("name", ["c1", "c1_neq"])
def test_error_bound_covers_table_perturbations(gpu_available, name):
    # Evaluation kind 3 bounds, per value, the change caused by table-value
    # errors up to (nu + nv) u sum|c_ab| plus the evaluation's own rounding.
    # Evaluations on tables perturbed inside that model (random and coherent
    # coefficient moves) and the FP64 matrix-core contraction must stay
    # within twice the bound on every mask.
    s = helpers.SETS[name]
    ctx = gpu_ctx(name)
    orders = [2, 4, 6, 8, 10, 11]
    configs = helpers.seeded_configs(5300 + len(name), orders)
    say(f"error bounds {name}")
    (bound,), ob, _ = run_batch(ctx, configs, "exact", kind=3)
    (ex,), _, _ = run_batch(ctx, configs, "exact")
    (mf,), _, _ = run_batch(ctx, configs, "mfma64")
    assert np.all(np.isfinite(bound)) and np.all(bound >= 0)
    u = 2.0 ** -53
    nonzero = ex != 0
    assert np.all(bound[nonzero] >= u * np.abs(ex[nonzero]))
    ratios = {"mfma64": float(np.max(np.abs(mf - ex) / np.maximum(bound, 1e-300)))}
    kappa = 2 * s["nc"]
    op = helpers.params_of(name)
    for label, seed, coherent in (("random", 1, False), ("coherent", 2, True)):
        pt = perturbed_tables(helpers.tables_of(name), kappa * u, seed, coherent)
        c2 = cdet_amd.Context(cdet_amd.make_params(s["N1"], s["N2"], s["nc"], s["beta"], s["U"], s["mu"],
                                                   s["alpha_up"], s["alpha_down"],
                                                   derived=(op.scale, op.diag_up_shift, op.diag_down_shift)))
        c2.upload_tables(pt)
        (pe,), _, _ = run_batch(c2, configs, "exact")
        c2.close()
        ratios[label] = float(np.max(np.abs(pe - ex) / np.maximum(bound, 1e-300)))
    say(f"{name}: max |difference| / bound: {ratios}")
    for k, r in ratios.items():
        assert r <= 2.0, (k, r)
    # the bound is not vacuous: the perturbed evaluations reach a visible
    # fraction of it somewhere
    assert max(ratios["random"], ratios["coherent"]) > 1e-4


def test_tuning_choices_bit_identical(gpu_available):
    # Every launch-shape choice of cdet_tuning gives the same values bit for
    # bit: contraction workgroups per CU and tile, paired classes, forked
    # streams, minors workgroup size per order, grouped-chain thresholds.
    import torch
    ctx = gpu_ctx("c1")
    default = ctx.tuning()
    orders = [2, 3, 5, 6, 8, 9, 10, 11, 12, 12] * 4 + [6] * 600
    configs = helpers.seeded_configs(9400, orders)
    choices = [
        {},
        {"contract_wg_per_cu": 4, "contract_tile": 64},
        {"contract_wg_per_cu": 5, "contract_tile": 16, "max_lanes": 1},
        {"pair_min_per_cu": 1, "max_lanes": 2},
        {"pair_min_per_cu": 1, "fused_minors_max_order": 10},
        {"pair_min_per_cu": 1 << 20},
        {"minors_threads": {11: 1024, 12: 1024}, "group16_min_terms": {n: 32 for n in range(15)},
         "group64_min_terms": {n: 64 for n in range(15)}},
        {"group16_min_terms": {n: 1 << 20 for n in range(15)}, "group64_min_terms": {n: 1 << 20 for n in range(15)}},
    ]
    ref = {}
    try:
        for mode in ("exact", "mfma64"):
            for k, ch in enumerate(choices):
                ctx.set_tuning(ch)
                say(f"tuning choice {k} {mode}")
                (out,), _, _ = run_batch(ctx, configs, mode)
                if k == 0:
                    ref[mode] = out
                else:
                    assert bits_equal(out, ref[mode]), (mode, ch)
        for bad in ({"contract_wg_per_cu": 9}, {"contract_tile": 32}, {"max_lanes": 5},
                    {"minors_threads": {12: 512}}, {"fused_minors_max_order": 11}):
            with pytest.raises(RuntimeError):
                ctx.set_tuning(bad)
    finally:
        ctx.set_tuning({})
    assert ctx.tuning() == default


def test_green_pair_reuse_bit_identical(gpu_available):
    # A plan with Green-pair reuse: the second run's configurations keep some
    # of the first run's vertices (at other positions) and draw the rest anew;
    # its values equal a fresh plan's and the oracle's bit for bit, the reused
    # pairs are counted, and after a table change nothing stale is reused.
    import torch
    from oracle import oracle as O
    name = "c1"
    ctx = gpu_ctx(name)
    orc = helpers.oracle_of(name)
    orders = [7, 5, 8, 6, 9, 4] * 12
    first = helpers.seeded_configs(9500, orders)
    fresh = helpers.seeded_configs(9501, orders)
    r = np.random.default_rng(95)
    second = []
    for a, b in zip(first, fresh):
        keep = r.random(len(a)) < 0.6
        v = np.where(keep[:, None], a, b)
        second.append(v[r.permutation(len(v))].copy())
    sptr = torch.cuda.current_stream().cuda_stream
    for graph in (False, True):
        plan = ctx.plan(orders, "mfma64" if graph else "exact")
        plan.reuse_pairs(True)
        out = torch.empty(plan.out_size, dtype=torch.float64, device="cuda")
        ctx.reset_counters()
        for cfgs in (first, second):
            d_v = torch.tensor(np.concatenate(cfgs), dtype=torch.float64, device="cuda")
            plan.run(d_v, out, None, use_graph=graph, stream=sptr)
            torch.cuda.synchronize()
        got = out.cpu().numpy()
        counters = ctx.counters()
        say(f"pair reuse (graph={graph}): {counters['green_pair_cache_hits']} reused, "
            f"{counters['green_pair_cache_misses']} looked up")
        assert counters["green_pair_cache_hits"] > 0
        (want,), ob, _ = run_batch(ctx, second, "mfma64" if graph else "exact")
        assert bits_equal(got, want)
        if not graph:
            for k in range(0, len(second), 7):
                assert bits_equal(got[ob[k]:ob[k + 1]], orc.all_subsets(second[k])), k
        plan.close()
    # table change between runs: the reusing plan reads the new tables only
    s = helpers.SETS[name]
    op = helpers.params_of(name)
    c2 = cdet_amd.Context(cdet_amd.make_params(s["N1"], s["N2"], s["nc"], s["beta"], s["U"], s["mu"],
                                               s["alpha_up"], s["alpha_down"],
                                               derived=(op.scale, op.diag_up_shift, op.diag_down_shift)))
    base = helpers.tables_of(name)
    c2.upload_tables(base)
    plan = c2.plan(orders, "exact")
    plan.reuse_pairs(True)
    out = torch.empty(plan.out_size, dtype=torch.float64, device="cuda")
    d_v = torch.tensor(np.concatenate(first), dtype=torch.float64, device="cuda")
    plan.run(d_v, out, None, use_graph=True, stream=sptr)
    torch.cuda.synchronize()
    made = {}
    bent = []
    for t in base:
        if id(t) not in made:
            made[id(t)] = O.Table(t.upper * (1.0 + 1e-3), t.lower * (1.0 - 1e-3), t.ucorner, t.lcorner)
        bent.append(made[id(t)])
    # the aliasing kinds take their own copies first (a kind another kind
    # aliases cannot be uploaded), then the sources, then the aliases again
    for k in (1, 3, 0, 2):
        c2.upload_table(k, bent[k].upper, bent[k].lower, bent[k].ucorner, bent[k].lcorner)
    c2.alias_table(1, 0)
    c2.alias_table(3, 2)
    plan.run(d_v, out, None, use_graph=True, stream=sptr)
    torch.cuda.synchronize()
    orc2 = O.Oracle(op, bent)
    got = out.cpu().numpy()
    _, ob = cdet_amd.packed_offsets(orders, False)
    for k in range(0, len(first), 5):
        assert bits_equal(got[ob[k]:ob[k + 1]], orc2.all_subsets(first[k])), k
    plan.close()
    c2.close()

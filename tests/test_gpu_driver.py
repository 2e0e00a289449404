"""MCMCMC driver on the GPU path vs the oracle: identical chosen subsets over
the 3x3 workload (1000 warmup + 1500 production steps, 2 chains) and
byte-identical checkpoints, including a 137 + 1363 step resume."""
import ctypes
import os

import numpy as np
import pytest

import cdet_amd
import helpers
import mcmcmc
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def c1_config(d, **over):
    c = mcmcmc.MCMCMCConfig()
    c.order_min, c.order_max, c.envelope_order, c.normalization_order = 1, 6, 7, 1
    for o in range(1, 7):
        c.lambdas[o] = 1.0
    c.lambda_warmup_steps, c.lambda_adapt_interval, c.lambda_adapt_rate = 1000, 200, 0.2
    c.production_steps, c.chain_count, c.base_seed = 1500, 2, 2026101401
    c.seed_mode = 0
    c.checkpoint_enabled, c.checkpoint_interval_steps, c.resume_mode = 1, 1000, 0
    c.checkpoint_dir = str(d).encode()
    for k, v in over.items():
        setattr(c, k, v)
    return c


def chk(d):
    return {n: open(os.path.join(d, n), "rb").read() for n in sorted(os.listdir(d)) if n.endswith(".chk")}


def gpu_ctx():
    s = helpers.SETS["c1"]
    op = helpers.params_of("c1")
    p = cdet_amd.make_params(s["N1"], s["N2"], s["nc"], s["beta"], s["U"], s["mu"], s["alpha_up"], s["alpha_down"],
                             derived=(op.scale, op.diag_up_shift, op.diag_down_shift))
    ctx = cdet_amd.Context(p)
    ctx.upload_tables(helpers.tables_of("c1"))
    return ctx


def test_driver_gpu_matches_oracle_and_resumes(gpu_available, tmp_path):
    ctx = gpu_ctx()
    gfn = ctypes.cast(cdet_amd.lib().cdet_eval_batch_host, ctypes.c_void_p)
    ofn = ctypes.cast(O.lib().orc_eval_batch_host, ctypes.c_void_p)
    orc = helpers.oracle_of("c1")
    print("driver gpu full", flush=True)
    rg, ratg, tg = mcmcmc.run(c1_config(tmp_path / "gpu"), gfn, ctx.handle, trace_steps=2500)
    print("driver oracle full", flush=True)
    ro, rato, to = mcmcmc.run(c1_config(tmp_path / "orc"), ofn, orc._h, trace_steps=2500)
    assert np.array_equal(tg, to)
    assert chk(tmp_path / "gpu") == chk(tmp_path / "orc")
    assert np.array_equal(ratg[1:], rato[1:], equal_nan=True)
    print("driver gpu resume", flush=True)
    mcmcmc.run(c1_config(tmp_path / "res", production_steps=137), gfn, ctx.handle)
    mcmcmc.run(c1_config(tmp_path / "res", resume_mode=2), gfn, ctx.handle)
    assert chk(tmp_path / "res") == chk(tmp_path / "gpu")
    for b in chk(tmp_path / "gpu").values():
        assert b"production_proposed_steps 1500\n" in b
    ctx.close()


def test_driver_mfma64_decisions(gpu_available, tmp_path):
    # FP64 matrix-core mode: report how many chosen subsets differ from the
    # bit-exact form (values differ at round-off only).
    ctx = gpu_ctx()
    gfn = ctypes.cast(cdet_amd.lib().cdet_eval_batch_host, ctypes.c_void_p)
    _, _, t_exact = mcmcmc.run(c1_config(tmp_path / "e", checkpoint_enabled=0), gfn, ctx.handle, trace_steps=2500)
    cdet_amd.lib().cdet_ctx_set_batch_mode(ctx.handle, cdet_amd.MODE_MFMA64)
    _, _, t_mfma = mcmcmc.run(c1_config(tmp_path / "m", checkpoint_enabled=0), gfn, ctx.handle, trace_steps=2500)
    first = np.argmax(t_exact != t_mfma, axis=1) if np.any(t_exact != t_mfma) else None
    print(f"mfma64 vs exact: differing steps {int(np.count_nonzero(t_exact != t_mfma))}, first {first}", flush=True)
    # chains agree at least until the first round-off flip
    assert np.array_equal(t_exact[:, :100], t_mfma[:, :100])
    ctx.close()


def test_driver_histogram_seed_gpu_matches_oracle(gpu_available, tmp_path):
    # Histogram seed: learning over 256 uniform envelopes (two batches of the
    # GPU path), then chains with the learned kernel.  Tables, chosen subsets
    # and checkpoints equal the oracle-driven run.
    ctx = gpu_ctx()
    gfn = ctypes.cast(cdet_amd.lib().cdet_eval_batch_host, ctypes.c_void_p)
    ofn = ctypes.cast(O.lib().orc_eval_batch_host, ctypes.c_void_p)
    orc = helpers.oracle_of("c1")

    def cfg(d, **over):
        return c1_config(d, seed_mode=1, N1=3, N2=3, histogram_learning_steps=256, histogram_tau_bins=8,
                         lambda_warmup_steps=400, production_steps=600, checkpoint_interval_steps=250, **over)

    print("histogram learn gpu / oracle", flush=True)
    hg, wg = mcmcmc.learn_histogram(cfg(tmp_path), gfn, ctx.handle)
    ho, wo = mcmcmc.learn_histogram(cfg(tmp_path), ofn, orc._h)
    assert np.array_equal(hg.tau, ho.tau) and np.array_equal(hg.space, ho.space)
    assert np.array_equal(hg.replica, ho.replica) and wg == wo
    print("histogram driver gpu", flush=True)
    _, ratg, tg = mcmcmc.run(cfg(tmp_path / "gpu"), gfn, ctx.handle, trace_steps=1000)
    print("histogram driver oracle", flush=True)
    _, rato, to = mcmcmc.run(cfg(tmp_path / "orc"), ofn, orc._h, trace_steps=1000)
    assert np.array_equal(tg, to)
    assert chk(tmp_path / "gpu") == chk(tmp_path / "orc")
    assert np.array_equal(ratg[1:], rato[1:], equal_nan=True)
    ctx.close()


def gpu_cond(ctx, seed, env_list):
    """cdet_seed_cond_batch_host over envelopes (list of (E, 4) arrays)."""
    orders = np.array([len(v) for v in env_list], dtype=np.int32)
    verts = np.ascontiguousarray(np.concatenate(env_list), dtype=np.float64)
    out = np.zeros(int(np.sum(1 << orders.astype(np.int64))))
    rc = cdet_amd.lib().cdet_seed_cond_batch_host(
        ctx.handle, ctypes.byref(seed), verts.ctypes.data, orders.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
        len(orders), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if rc != 0:
        raise RuntimeError(cdet_amd.lib().cdet_last_error().decode())
    return orders, out


def test_gpu_conditional_probabilities_bit_identical(gpu_available):
    # the GPU recursion equals the driver's host recursion to the last bit,
    # envelopes 2..14, on learned-like random tables
    ctx = gpu_ctx()
    r = np.random.default_rng(5)
    bins, n1, n2 = 8, 3, 3

    def rows(k):
        m = r.uniform(0.02, 1.0, (k, k))
        return m / m.sum(axis=1, keepdims=True)

    seed = mcmcmc.HistogramSeed.make(bins, n1, n2, rows(bins), rows(n1 * n2), rows(2))
    envs = [r.uniform(0.0, 1.0, (E, 4)) for E in list(range(2, 15)) + [14, 9, 3]]
    print("gpu conditional probabilities", flush=True)
    orders, out = gpu_cond(ctx, seed, envs)
    off = 0
    for E, v in zip(orders, envs):
        want = mcmcmc.conditional_probabilities(seed, v)
        got = out[off:off + (1 << E)]
        assert np.array_equal(got, want), E
        off += 1 << E
    # a negative table entry is reported with the reference's message
    bad = mcmcmc.HistogramSeed.make(bins, n1, n2, -rows(bins), rows(n1 * n2), rows(2))
    with pytest.raises(RuntimeError, match="MCMCMC seed density is negative or non-finite"):
        gpu_cond(ctx, bad, envs[:2])
    ctx.close()


def test_driver_histogram_gpu_conditional_backend(gpu_available, tmp_path):
    # chains with the GPU conditional-generation callback equal those with the
    # driver's host recursion (per-chain envelopes, checkpoints included)
    ctx = gpu_ctx()
    gfn = ctypes.cast(cdet_amd.lib().cdet_eval_batch_host, ctypes.c_void_p)
    cfn = ctypes.cast(cdet_amd.lib().cdet_seed_cond_batch_host, ctypes.c_void_p)

    def cfg(d):
        return c1_config(d, seed_mode=1, N1=3, N2=3, histogram_learning_steps=128, histogram_tau_bins=8,
                         lambda_warmup_steps=300, production_steps=400, checkpoint_interval_steps=200,
                         chain_count=4)

    h, _ = mcmcmc.learn_histogram(cfg(tmp_path), gfn, ctx.handle)
    env = [7, 5, 10, 3]
    print("histogram driver, host recursion", flush=True)
    _, _, t_host = mcmcmc.run_backend(cfg(tmp_path / "h"), gfn, ctx.handle, seed=h, envelopes=env, trace_steps=700)
    print("histogram driver, GPU recursion", flush=True)
    _, _, t_gpu = mcmcmc.run_backend(cfg(tmp_path / "g"), gfn, ctx.handle, cond_fn=cfn, cond_user=ctx.handle,
                                     seed=h, envelopes=env, trace_steps=700)
    assert np.array_equal(t_host, t_gpu)
    assert chk(tmp_path / "h") == chk(tmp_path / "g")
    ctx.close()


def test_driver_two_groups_in_flight_gpu(gpu_available, tmp_path):
    # cdet_eval_batch_submit / wait: two groups of chains in flight on the
    # two slots' streams take exactly the steps of the synchronous host
    # callback (5 chains: groups of 3 and 2), in both batch modes.
    ctx = gpu_ctx()
    L = cdet_amd.lib()
    hfn = ctypes.cast(L.cdet_eval_batch_host, ctypes.c_void_p)
    sfn = ctypes.cast(L.cdet_eval_batch_submit, ctypes.c_void_p)
    wfn = ctypes.cast(L.cdet_eval_batch_wait, ctypes.c_void_p)
    env = [7, 5, 6, 3, 7]
    for mode in ("exact", "mfma64"):
        assert L.cdet_ctx_set_batch_mode(ctx.handle, cdet_amd.MODES[mode]) == 0
        c = c1_config(tmp_path / mode, chain_count=5, checkpoint_enabled=0, lambda_warmup_steps=300,
                      production_steps=500)
        print(f"driver two groups {mode} sync", flush=True)
        rs, rats, ts = mcmcmc.run_backend(c, hfn, ctx.handle, envelopes=env, trace_steps=800)
        print(f"driver two groups {mode} async", flush=True)
        try:
            ra, rata, ta = mcmcmc.run_backend(c, hfn, ctx.handle, envelopes=env, trace_steps=800, submit_fn=sfn,
                                              wait_fn=wfn)
        except RuntimeError as e:
            raise RuntimeError(f"{e}; library: {L.cdet_last_error().decode()}") from e
        assert np.array_equal(ts, ta)
        assert np.array_equal(rats, rata, equal_nan=True)
        for a, b in zip(rs, ra):
            assert list(a.production_chosen_counts) == list(b.production_chosen_counts)
    # a slot holds one batch: a second submit before its wait is refused
    v = np.zeros((2, 4))
    v[:, 0] = [0.1, 0.2]
    o = np.array([2], dtype=np.int32)
    vp = v.ctypes.data_as(ctypes.c_void_p)
    op = o.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    assert L.cdet_eval_batch_submit(ctx.handle, 0, vp, op, 1) == 0
    assert L.cdet_eval_batch_submit(ctx.handle, 0, vp, op, 1) != 0
    out = np.zeros(4)
    assert L.cdet_eval_batch_wait(ctx.handle, 0, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    assert L.cdet_eval_batch_wait(ctx.handle, 0, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) != 0
    ref = np.zeros(4)
    assert L.cdet_eval_batch_host(ctx.handle, vp, op, 1, ref.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    assert np.array_equal(out, ref)
    ctx.close()


def test_driver_pinned_value_buffers_gpu(gpu_available, tmp_path):
    # Value buffers from cdet_host_alloc (pinned, filled by DMA with no
    # staging copy) give the steps and statistics of the driver's own
    # buffers, synchronously and with two groups in flight; a buffer inside
    # an allocation, or a pageable one, still gets the values.
    ctx = gpu_ctx()
    L = cdet_amd.lib()
    hfn = ctypes.cast(L.cdet_eval_batch_host, ctypes.c_void_p)
    sfn = ctypes.cast(L.cdet_eval_batch_submit, ctypes.c_void_p)
    wfn = ctypes.cast(L.cdet_eval_batch_wait, ctypes.c_void_p)
    ha = ctypes.cast(L.cdet_host_alloc, ctypes.c_void_p)
    hf = ctypes.cast(L.cdet_host_free, ctypes.c_void_p)
    env = [7, 5, 6, 3, 7]
    assert L.cdet_ctx_set_batch_mode(ctx.handle, cdet_amd.MODES["mfma64"]) == 0
    c = c1_config(tmp_path, chain_count=5, checkpoint_enabled=0, lambda_warmup_steps=200, production_steps=300)
    print("driver pinned buffers: vector buffers", flush=True)
    r0, rat0, t0 = mcmcmc.run_backend(c, hfn, ctx.handle, envelopes=env, trace_steps=500)
    for name, kw in (("sync", {}), ("async", {"submit_fn": sfn, "wait_fn": wfn})):
        print(f"driver pinned buffers: {name}", flush=True)
        r1, rat1, t1 = mcmcmc.run_backend(c, hfn, ctx.handle, envelopes=env, trace_steps=500, host_alloc_fn=ha,
                                          host_free_fn=hf, **kw)
        assert np.array_equal(t0, t1), name
        assert np.array_equal(rat0, rat1, equal_nan=True), name
        for a, b in zip(r0, r1):
            assert list(a.order_signal_mean) == list(b.order_signal_mean), name
    v = np.array(helpers.seeded_configs(4321, [6])[0])
    o = np.array([6], dtype=np.int32)
    vp = v.ctypes.data_as(ctypes.c_void_p)
    op = o.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    ref = np.zeros(64)
    assert L.cdet_eval_batch_host(ctx.handle, vp, op, 1, ref.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    buf = L.cdet_host_alloc(ctx.handle, 8 * 80)
    assert buf
    inner = (ctypes.c_double * 64).from_address(buf + 8 * 16)
    assert L.cdet_eval_batch_host(ctx.handle, vp, op, 1, inner) == 0
    assert np.array_equal(np.ctypeslib.as_array(inner), ref)
    L.cdet_host_free(ctx.handle, buf)
    ctx.close()


def test_driver_two_groups_forked_slot_plans(gpu_available, tmp_path):
    # Envelopes up to 12: each slot's plan spreads its order buckets over two
    # streams and launches from the slot's own thread; the chains still take
    # exactly the steps of the synchronous host callback.
    ctx = gpu_ctx()
    L = cdet_amd.lib()
    hfn = ctypes.cast(L.cdet_eval_batch_host, ctypes.c_void_p)
    sfn = ctypes.cast(L.cdet_eval_batch_submit, ctypes.c_void_p)
    wfn = ctypes.cast(L.cdet_eval_batch_wait, ctypes.c_void_p)
    env = [12, 11, 10, 9, 12, 11, 10, 8]
    c = c1_config(tmp_path, chain_count=len(env), checkpoint_enabled=0, lambda_warmup_steps=100,
                  production_steps=200)
    print("forked slots: sync", flush=True)
    rs, rats, ts = mcmcmc.run_backend(c, hfn, ctx.handle, envelopes=env, trace_steps=300)
    print("forked slots: async", flush=True)
    try:
        ra, rata, ta = mcmcmc.run_backend(c, hfn, ctx.handle, envelopes=env, trace_steps=300, submit_fn=sfn,
                                          wait_fn=wfn)
    except RuntimeError as e:
        raise RuntimeError(f"{e}; library: {L.cdet_last_error().decode()}") from e
    assert np.array_equal(ts, ta)
    assert np.array_equal(rats, rata, equal_nan=True)
    ctx.close()


def test_c4_sweeps_mfma64_agree_with_fp64_within_two_sigma(gpu_available):
    # The C4 pass criterion: 10^4 lockstep sweeps of the 256 chains of the 6x6
    # workload (envelopes max(3, 2 + floor(11 u^2))) with the FP64 matrix-core
    # contraction and with the bit-identical FP64 form; per order, the
    # cross-chain means of the chains' order-signal means (the reference's CLT
    # samples, include/cdet/mcmcmc_sampler.h:1117-1131) agree within 2 sigma.
    import bench
    import workloads
    w = workloads.C4
    orders, _ = workloads.configs("c4")
    print("c4: building tables", flush=True)
    ctx = workloads.context("c4")
    stats = {}
    for mode in ("mfma64", "exact"):
        print(f"c4: 10^4 sweeps, {mode}", flush=True)
        cdet_amd.lib().cdet_ctx_set_batch_mode(ctx.handle, cdet_amd.MODES[mode])
        bench.run_sweeps(ctx, w, orders, 10000, w["envelope"] - 1)
        stats[mode] = bench.run_sweeps.last_stats
    ag = mcmcmc.agree_within(stats["mfma64"], stats["exact"], 2.0)
    print("c4: z per order", {o: round(z, 3) for o, (z, _) in ag.items()}, flush=True)
    assert all(a for _, a in ag.values()), ag
    # the even orders carry signal (odd orders vanish at half filling)
    assert stats["exact"][2][0] != 0.0 and stats["exact"][4][0] != 0.0
    ctx.close()

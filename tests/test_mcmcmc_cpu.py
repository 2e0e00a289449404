"""Driver checks on the CPU: the lockstep MCMCMC driver over the oracle and
over the reference's synthetic product integrand."""
import ctypes
import os

import numpy as np
import pytest

import helpers
import mcmcmc
from oracle import oracle as O


def base_config(tmp, **over):
    c = mcmcmc.MCMCMCConfig()
    c.order_min, c.order_max, c.envelope_order, c.normalization_order = 1, 3, 4, 1
    for o in (1, 2, 3):
        c.lambdas[o] = 1.0
    c.lambda_warmup_steps, c.lambda_adapt_interval, c.lambda_adapt_rate = 200, 60, 0.2
    c.production_steps = 300
    c.chain_count = 2
    c.base_seed = 2026101401
    c.seed_mode = 0
    c.checkpoint_enabled = 1
    c.checkpoint_interval_steps = 100
    c.resume_mode = 0
    c.checkpoint_dir = str(tmp).encode()
    for k, v in over.items():
        setattr(c, k, v.encode() if k == "checkpoint_dir" else v)
    return c


def oracle_eval():
    orc = helpers.oracle_of("ref2x2")
    return ctypes.cast(O.lib().orc_eval_batch_host, ctypes.c_void_p), orc


def chk_bytes(d):
    return {n: open(os.path.join(d, n), "rb").read() for n in sorted(os.listdir(d)) if n.endswith(".chk")}


def test_synthetic_integral_ratios_recovered():
    # tests/test_mcmcmc_synthetic.cpp:135-159: orders 1..4, envelope 5,
    # product integrand with integral 1 at every order.
    c = mcmcmc.MCMCMCConfig()
    c.order_min, c.order_max, c.envelope_order, c.normalization_order = 1, 4, 5, 1
    for o, l in zip((1, 2, 3, 4), (0.7, 1.1, 1.6, 2.0)):
        c.lambdas[o] = l
    c.lambda_warmup_steps, c.lambda_adapt_interval, c.lambda_adapt_rate = 20000, 1000, 0.25
    c.production_steps, c.chain_count, c.base_seed = 400000, 1, 0x7812BC91
    fn = ctypes.cast(O.lib().orc_synthetic_batch_eval, ctypes.c_void_p)
    res, ratios, _ = mcmcmc.run(c, fn, None)
    for o in (2, 3, 4):
        assert abs(ratios[o] - 1.0) < 0.04, ratios
        assert res[0].production_chosen_counts[o] > 100


def test_checkpoint_resume_byte_identical(tmp_path):
    # tests/test_checkpoint_resume.py: full run vs shorter run + resume with
    # MCMCMC_RESUME=require; final checkpoints byte-identical.
    fn, orc = oracle_eval()
    full = tmp_path / "full"
    part = tmp_path / "part"
    mcmcmc.run(base_config(full), fn, orc._h)
    mcmcmc.run(base_config(part, production_steps=137), fn, orc._h)
    res, _, _ = mcmcmc.run(base_config(part, resume_mode=2), fn, orc._h)
    a, b = chk_bytes(full), chk_bytes(part)
    assert len(a) == len(b) == 2
    for k in a:
        assert a[k] == b[k], k
        assert b"production_proposed_steps 300\n" in a[k]
    assert res[0].production_proposed_steps == 300


def test_interrupted_run_resumes_identically(tmp_path):
    fn, orc = oracle_eval()
    full = tmp_path / "full"
    cut = tmp_path / "cut"
    mcmcmc.run(base_config(full), fn, orc._h)
    # killed after 333 steps: last periodic checkpoint at 300 completed steps
    mcmcmc.run(base_config(cut, stop_after_steps=333), fn, orc._h)
    mcmcmc.run(base_config(cut, resume_mode=2), fn, orc._h)
    assert chk_bytes(full) == chk_bytes(cut)


def test_resume_require_without_checkpoint_fails(tmp_path):
    fn, orc = oracle_eval()
    with pytest.raises(RuntimeError, match="checkpoint is missing"):
        mcmcmc.run(base_config(tmp_path / "none", resume_mode=2), fn, orc._h)


def test_lockstep_chains_are_independent(tmp_path):
    fn, orc = oracle_eval()
    r3, _, t3 = mcmcmc.run(base_config(tmp_path / "a", chain_count=3, checkpoint_enabled=0), fn, orc._h,
                           trace_steps=500)
    r1, _, t1 = mcmcmc.run(base_config(tmp_path / "b", chain_count=1, checkpoint_enabled=0), fn, orc._h,
                           trace_steps=500)
    assert np.array_equal(t3[0], t1[0])
    assert list(r3[0].order_signal_mean) == list(r1[0].order_signal_mean)


def test_params_file_parsing(tmp_path):
    p = tmp_path / "params.txt"
    p.write_text("# comment\nN1=3\nN2=3\nU=2.0\nBETA=4\nALPHA_UP=1\nALPHA_DOWN=1\n"
                 "MCMCMC_ORDER_MIN=1\nMCMCMC_ORDER_MAX=6\nMCMCMC_ENVELOPE_ORDER=7\nMCMCMC_LAMBDAS=1\n"
                 "MCMCMC_CHAIN_COUNT=2\nMCMCMC_BASE_SEED=2026101401\nMCMCMC_RESUME=required\n")
    params = mcmcmc.read_params(str(p))
    c = mcmcmc.config_from_params(params)
    assert (c.order_min, c.order_max, c.envelope_order, c.normalization_order) == (1, 6, 7, 1)
    assert list(c.lambdas[1:7]) == [1.0] * 6 and c.resume_mode == 2 and c.chain_count == 2
    assert mcmcmc.integrand_scalars_from_params(params) == (72.0 ** 0.5, 0.5, 0.5)


def test_per_chain_envelopes_match_single_envelope_runs(tmp_path):
    # chain k with envelope E_k is the reference sampler at envelope_order =
    # E_k and order_max = min(order_max, E_k - 1), seeded by its index
    fn, orc = oracle_eval()
    env = [4, 3, 5]
    c = base_config(tmp_path / "mix", chain_count=3, checkpoint_enabled=0)
    c.envelope_order = 5
    _, _, t_mix = mcmcmc.run_envelopes(c, env, fn, orc._h, trace_steps=500)
    for k, e in enumerate(env):
        one = base_config(tmp_path / f"one{k}", chain_count=k + 1, checkpoint_enabled=0)
        one.envelope_order, one.order_max = e, min(3, e - 1)
        _, _, t_one = mcmcmc.run(one, fn, orc._h, trace_steps=500)
        assert np.array_equal(t_mix[k], t_one[k]), k
    # equal envelopes reproduce the shared-envelope run
    same = base_config(tmp_path / "s", chain_count=2, checkpoint_enabled=0)
    _, _, t_same = mcmcmc.run_envelopes(same, [4, 4], fn, orc._h, trace_steps=500)
    _, _, t_ref = mcmcmc.run(base_config(tmp_path / "r", chain_count=2, checkpoint_enabled=0), fn, orc._h,
                             trace_steps=500)
    assert np.array_equal(t_same, t_ref)


def test_per_chain_envelopes_resume_and_ratios(tmp_path):
    fn, orc = oracle_eval()
    env = [5, 3, 4, 2]
    full, cut = tmp_path / "full", tmp_path / "cut"
    mk = lambda d, **o: base_config(d, chain_count=4, **o)
    mcmcmc.run_envelopes(mk(full), env, fn, orc._h)
    mcmcmc.run_envelopes(mk(cut, stop_after_steps=333), env, fn, orc._h)
    mcmcmc.run_envelopes(mk(cut, resume_mode=2), env, fn, orc._h)
    a, b = chk_bytes(full), chk_bytes(cut)
    assert a == b and len(a) == 4
    assert b"envelope_order 2\norder_min 1\n" not in a["chain_000003.chk"]
    assert b"order_max 1\nenvelope_order 2\n" in a["chain_000003.chk"]
    # synthetic product integrand: every order integrates to one
    s = mcmcmc.MCMCMCConfig()
    s.order_min, s.order_max, s.envelope_order, s.normalization_order = 1, 4, 5, 1
    for o in (1, 2, 3, 4):
        s.lambdas[o] = 1.0
    s.lambda_warmup_steps, s.lambda_adapt_interval, s.lambda_adapt_rate = 20000, 1000, 0.25
    s.production_steps, s.chain_count, s.base_seed = 200000, 4, 0x7812BC91
    res, ratios, _ = mcmcmc.run_envelopes(s, [5, 5, 4, 3], ctypes.cast(O.lib().orc_synthetic_batch_eval,
                                                                     ctypes.c_void_p), None)
    assert all(abs(ratios[o] - 1.0) < 0.05 for o in (2, 3, 4)), ratios
    assert res[3].order_signal_count[3] == 0 and res[0].order_signal_count[4] == 200000


def test_two_groups_in_flight_take_the_lockstep_steps(tmp_path):
    # An asynchronous backend (submit + wait) puts the two halves of the
    # chains in flight in turn; each chain takes exactly its lockstep steps.
    # The submit/wait pair here wraps the oracle's batch callback and records
    # the calls, so the test also checks that the slots alternate and every
    # submitted batch is waited for once.
    fn, orc = oracle_eval()
    eval_c = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                              ctypes.c_int64, ctypes.POINTER(ctypes.c_double))(fn.value)
    pending, calls = {}, []

    def submit(user, slot, verts, orders, count):
        n = [orders[i] for i in range(count)]
        nv = sum(n)
        v = np.ctypeslib.as_array(ctypes.cast(verts, ctypes.POINTER(ctypes.c_double)), shape=(4 * nv,)).copy()
        if slot in pending:
            return 7
        pending[slot] = (v, np.array(n, dtype=np.int32))
        calls.append(("submit", slot, count))
        return 0

    def wait(user, slot, out):
        if slot not in pending:
            return 8
        v, n = pending.pop(slot)
        calls.append(("wait", slot))
        return eval_c(orc._h, v.ctypes.data_as(ctypes.c_void_p), n.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                      len(n), out)

    sfn, wfn = mcmcmc.SUBMIT_FN(submit), mcmcmc.WAIT_FN(wait)
    env = [4, 3, 4, 2, 3]
    c = base_config(tmp_path / "a", chain_count=5, checkpoint_enabled=0)
    c.envelope_order = 4
    c.lambda_warmup_steps, c.production_steps = 60, 90
    r_sync, _, t_sync = mcmcmc.run_backend(c, fn, orc._h, envelopes=env, trace_steps=150)
    r_async, _, t_async = mcmcmc.run_backend(c, fn, orc._h, envelopes=env, trace_steps=150, submit_fn=sfn,
                                             wait_fn=wfn)
    assert np.array_equal(t_sync, t_async)
    for a, b in zip(r_sync, r_async):
        assert list(a.order_signal_mean) == list(b.order_signal_mean)
        assert list(a.production_chosen_counts) == list(b.production_chosen_counts)
    assert not pending
    subs = [x for x in calls if x[0] == "submit"]
    assert {s[1] for s in subs} == {0, 1} and sum(1 for x in calls if x[0] == "wait") == len(subs)
    assert {s[2] for s in subs if s[1] == 0} == {3} and {s[2] for s in subs if s[1] == 1} == {2}


def test_table_backend_and_workflow_resolution():
    # src/main.cpp:322-345, :428-446: names are normalised, aliases mapped,
    # environment variables override the file
    base = {"N1": ["3"], "N2": ["3"], "PBC": ["true"], "SUBREGION_MASK": ["first_row"]}
    wf, be, cd = mcmcmc.table_workflow_from_params(base, environ={})
    assert (wf, be, cd) == ("resident", "resident", "chunked_green_tables")
    p = dict(base, LOAD_DATA_WORKFLOW=["Load-Prebuilt"], MCMCMC_TABLE_BACKEND=["LRU"],
             LOAD_DATA_CHUNK_DIR=["tables"])
    assert mcmcmc.table_workflow_from_params(p, environ={}) == ("load_chunks", "disk_lru", "tables")
    for alias in ("on_demand", "disk", "disk-cache", "DISK_LRU"):
        p = dict(base, MCMCMC_TABLE_BACKEND=[alias])
        assert mcmcmc.table_workflow_from_params(p, environ={})[1] == "disk_lru"
    env = {"CDET_MCMCMC_TABLE_BACKEND": "disk_lru", "CDET_LOAD_DATA_WORKFLOW": "load_chunks",
           "CDET_LOAD_DATA_CHUNK_DIR": "/x"}
    assert mcmcmc.table_workflow_from_params(base, environ=env) == ("load_chunks", "disk_lru", "/x")
    with pytest.raises(RuntimeError, match="MCMCMC_TABLE_BACKEND must be resident or disk_lru"):
        mcmcmc.table_workflow_from_params(dict(base, MCMCMC_TABLE_BACKEND=["mmap"]), environ={})


def test_load_data_dispatch_without_gpu():
    # load_data() (include/cdet/load_data.h:119-178): load_chunks goes to the
    # chunk loader for both backends (disk_lru maps to HBM residency), the
    # resident workflow to the table builder; the reference's errors otherwise
    calls = []

    class Ctx:
        def load_chunks(self, d):
            calls.append(("load", d))

        def build_tables(self, **kw):
            calls.append(("build", kw))

    base = {"N1": ["3"], "N2": ["3"], "PBC": ["true"], "SUBREGION_MASK": ["origin_square"],
            "SUBREGION_SQUARE_RADIUS": ["1"]}
    for backend in ("resident", "disk_lru"):
        p = dict(base, LOAD_DATA_WORKFLOW=["load_chunks"], MCMCMC_TABLE_BACKEND=[backend],
                 LOAD_DATA_CHUNK_DIR=["d"], MCMCMC_LRU_CACHE_GB=["0.1"])
        r = mcmcmc.load_data_from_params(Ctx(), p, environ={})
        assert r == {"workflow": "load_chunks", "backend": backend, "residency": "hbm"}
    assert calls == [("load", "d"), ("load", "d")]
    r = mcmcmc.load_data_from_params(Ctx(), base, environ={})
    assert calls[-1] == ("build", {"pbc": True, "mask": "origin_square", "mask_param": 1})
    with pytest.raises(RuntimeError, match="requires LOAD_DATA_WORKFLOW=load_chunks"):
        mcmcmc.load_data_from_params(Ctx(), dict(base, MCMCMC_TABLE_BACKEND=["disk_lru"]), environ={})
    with pytest.raises(RuntimeError, match="can only run with LOAD_DATA_WORKFLOW=resident or load_chunks"):
        mcmcmc.load_data_from_params(Ctx(), dict(base, LOAD_DATA_WORKFLOW=["sample"]), environ={})


def test_cross_chain_stats_are_chain_means_as_clt_samples():
    # include/cdet/mcmcmc_sampler.h:1117-1131: one sample per chain with
    # signal at the order; running mean and M2 as RunningStatsLongDouble
    r = np.random.default_rng(3)
    res = (mcmcmc.ChainResult * 5)()
    means = r.normal(size=(5, 4))
    for c in range(5):
        for o in range(2, 6):
            res[c].order_signal_mean[o] = means[c, o - 2]
            res[c].order_signal_count[o] = 0 if (c == 4 and o == 5) else 10
    st = mcmcmc.cross_chain_stats(list(res), 2, 5)
    for o in range(2, 6):
        x = means[:, o - 2] if o != 5 else means[:4, 3]
        m, se, n = st[o]
        assert n == len(x) and abs(m - x.mean()) < 1e-14
        assert abs(se - x.std(ddof=1) / np.sqrt(len(x))) < 1e-14
    ag = mcmcmc.agree_within(st, {o: (st[o][0] + 1.5 * st[o][1], st[o][1], st[o][2]) for o in st})
    assert all(a for _, a in ag.values())  # 1.5 se apart, combined se = sqrt(2) se: z ~ 1.06
    ag = mcmcmc.agree_within(st, {o: (st[o][0] + 3.0 * st[o][1], st[o][1], st[o][2]) for o in st})
    assert not any(a for _, a in ag.values())  # z ~ 2.12

"""Histogram seed of the MCMCMC driver on the CPU (include/cdet/mcmcmc_seed.h,
include/cdet/mcmcmc_sampler.h:252-330, :607-671).

Pins: the reference's own synthetic test with its positive histogram seed
(tests/test_mcmcmc_synthetic.cpp:98-180); the subset recursion of the
conditional generation probabilities against a brute-force average over
generation orders; closed forms (uniform tables give probability one); the
sampled child distribution against the table rows; the learned tables
against an independent float64 restatement of the pair accumulation on the
same random stream; the export format; resume and thread independence."""
import ctypes
import itertools
import math
import os

import numpy as np
import pytest

import helpers
import mcmcmc
import rng
from oracle import oracle as O


def synthetic_fn():
    return ctypes.cast(O.lib().orc_synthetic_batch_eval, ctypes.c_void_p)


def positive_seed():
    # tests/test_mcmcmc_synthetic.cpp:98-117
    tau = np.full((4, 4), 0.1)
    space = np.full((4, 4), 0.1)
    np.fill_diagonal(tau, 0.7)
    np.fill_diagonal(space, 0.7)
    return mcmcmc.HistogramSeed.make(4, 2, 2, tau, space, [0.8, 0.2, 0.2, 0.8], uniform_mix=0.2)


def random_seed(bins, n1, n2, seed):
    r = np.random.default_rng(seed)
    S = n1 * n2

    def rows(k):
        m = r.uniform(0.05, 1.0, (k, k))
        return m / m.sum(axis=1, keepdims=True)

    return mcmcmc.HistogramSeed.make(bins, n1, n2, rows(bins), rows(S), rows(2))


def py_bin(v, n):
    return min(max(int(math.floor(v * n)), 0), n - 1)


def py_density(h, c, p):
    S = h.N1 * h.N2
    site = lambda v: py_bin(v[1], h.N1) * h.N2 + py_bin(v[2], h.N2)
    return (h.tau[py_bin(p[0], h.tau_bins), py_bin(c[0], h.tau_bins)] * h.tau_bins
            * h.space[site(p), site(c)] * S
            * h.replica[py_bin(p[3], 2), py_bin(c[3], 2)] * 2.0)


def brute_conditional(h, verts):
    """Average over generation orders of the sequential regeneration density:
    for generated set G (|G| = m) and order c_1..c_m, vertex c_j is the child
    of a uniformly chosen vertex among the kept ones and c_1..c_{j-1}."""
    E = len(verts)
    out = np.zeros(1 << E)
    out[0] = 1.0
    q = lambda c, p: py_density(h, verts[c], verts[p])
    for G in range(1, (1 << E) - 1):
        members = [i for i in range(E) if G >> i & 1]
        kept = [i for i in range(E) if not G >> i & 1]
        tot = 0.0
        for perm in itertools.permutations(members):
            prod = 1.0
            for j, c in enumerate(perm):
                parents = kept + list(perm[:j])
                prod *= sum(q(c, p) for p in parents) / len(parents)
            tot += prod
        out[G] = tot / math.factorial(len(members))
    return out


def test_reference_positive_seed_recovers_equal_ratios():
    # tests/test_mcmcmc_synthetic.cpp:132-180: orders 1..4, envelope 5,
    # product integrand with integral 1 at every order; histogram seed
    # tolerance 0.07 over 400000 production steps.
    c = mcmcmc.MCMCMCConfig(N1=2, N2=2)
    c.order_min, c.order_max, c.envelope_order, c.normalization_order = 1, 4, 5, 1
    for o, l in zip((1, 2, 3, 4), (0.7, 1.1, 1.6, 2.0)):
        c.lambdas[o] = l
    c.lambda_warmup_steps, c.lambda_adapt_interval, c.lambda_adapt_rate = 20000, 1000, 0.25
    c.production_steps, c.chain_count, c.base_seed = 400000, 1, 0xD9024AD7
    c.seed_mode = 1
    c.histogram_learning_steps = 1
    res, ratios, _ = mcmcmc.run_seeded(c, positive_seed(), synthetic_fn(), None)
    for o in (2, 3, 4):
        assert abs(ratios[o] - 1.0) < 0.07, ratios
        assert res[0].production_chosen_counts[o] > 100


def test_positive_seed_keeps_support():
    # tests/test_mcmcmc_synthetic.cpp:162-174
    probe = np.array([[x] * 4 for x in (0.01, 0.29, 0.51, 0.79, 0.99)])
    p = mcmcmc.conditional_probabilities(positive_seed(), probe)
    assert np.all(np.isfinite(p[1:-1])) and np.all(p[1:-1] > 0.0)
    assert p[0] == 1.0 and p[-1] == 0.0


@pytest.mark.parametrize("E,dims", [(4, (3, 2, 3)), (5, (4, 2, 2))])
def test_conditional_recursion_matches_generation_orders(E, dims):
    h = random_seed(*dims, seed=E)
    verts = np.random.default_rng(100 + E).uniform(0.0, 1.0, (E, 4))
    got = mcmcmc.conditional_probabilities(h, verts)
    want = brute_conditional(h, verts)
    np.testing.assert_allclose(got[:-1], want[:-1], rtol=1e-12, atol=0.0)
    assert got[-1] == 0.0


def test_uniform_tables_give_probability_one():
    bins, n1, n2 = 5, 2, 3
    S = n1 * n2
    h = mcmcmc.HistogramSeed.make(bins, n1, n2, np.full((bins, bins), 1.0 / bins), np.full((S, S), 1.0 / S),
                                  [0.5] * 4)
    verts = np.random.default_rng(7).uniform(0.0, 1.0, (6, 4))
    p = mcmcmc.conditional_probabilities(h, verts)
    np.testing.assert_allclose(p[:-1], 1.0, rtol=1e-13)
    assert p[-1] == 0.0
    # the uniform seed itself: all ones, the full set included (:259-262)
    assert np.all(mcmcmc.conditional_probabilities(None, verts) == 1.0)


def test_density_indexing():
    h = random_seed(4, 2, 3, seed=11)
    r = np.random.default_rng(12)
    for _ in range(50):
        c, p = r.uniform(0.0, 1.0, 4), r.uniform(0.0, 1.0, 4)
        assert mcmcmc.seed_density(h, c, p) == pytest.approx(py_density(h, c, p), rel=1e-15)
    assert mcmcmc.seed_density(None, c, p) == 1.0


def test_children_follow_table_rows():
    bins, n1, n2 = 4, 2, 3
    h = random_seed(bins, n1, n2, seed=21)
    parent = np.array([0.6, 0.9, 0.4, 0.7])   # tau bin 2, site x=1,y=1 -> 4, replica 1
    n = 200000
    kids = mcmcmc.sample_children(h, np.tile(parent, (n, 1)), 2026)
    assert np.all((kids >= 0.0) & (kids < 1.0))
    tb = np.floor(kids[:, 0] * bins).astype(int)
    xs = np.floor(kids[:, 1] * n1).astype(int)
    ys = np.floor(kids[:, 2] * n2).astype(int)
    rb = np.floor(kids[:, 3] * 2).astype(int)
    for got_idx, row in ((tb, h.tau[2]), (xs * n2 + ys, h.space[1 * n2 + 1]), (rb, h.replica[1])):
        freq = np.bincount(got_idx, minlength=len(row)) / n
        sd = np.sqrt(row * (1 - row) / n)
        assert np.all(np.abs(freq - row) < 5 * sd + 1e-12), (freq, row)
    # the stream is a pure function of the seed
    again = mcmcmc.sample_children(h, np.tile(parent, (1000, 1)), 2026)
    assert np.array_equal(again, kids[:1000])


def py_learn(c, bins, n1, n2, mix, pseudo):
    """float64 restatement of the learning pass over the same stream."""
    E, L = c.envelope_order, c.histogram_learning_steps
    seed = rng.derive_mcmcmc_seed(c.base_seed, 0x484953544F534545, 0)
    v = O.uniform_stream(seed, 0.0, 1.0, 4 * E * L).reshape(L, E, 4)
    S = n1 * n2
    T, X, R = np.zeros((bins, bins)), np.zeros((S, S)), np.zeros((2, 2))
    site = lambda u: py_bin(u[1], n1) * n2 + py_bin(u[2], n2)
    for env in v:
        for mask in range(1, 1 << E):
            members = [i for i in range(E) if mask >> i & 1]
            o = len(members)
            if o < max(2, c.order_min) or o > c.order_max:
                continue
            val = np.prod([0.5 + env[i][0] for i in members])
            w = abs(val) / math.comb(E, o) / (o * (o - 1))
            for p in members:
                for ch in members:
                    if ch == p:
                        continue
                    T[py_bin(env[p][0], bins), py_bin(env[ch][0], bins)] += w
                    X[site(env[p]), site(env[ch])] += w
                    R[py_bin(env[p][3], 2), py_bin(env[ch][3], 2)] += w

    def norm(M):
        k = M.shape[1]
        den = M.sum(axis=1, keepdims=True) + pseudo * k
        learned = np.where(den > 0, (M + pseudo) / np.where(den > 0, den, 1), 1.0 / k)
        P = (1 - mix) * learned + mix / k
        return P / P.sum(axis=1, keepdims=True)

    return norm(T), norm(X), norm(R)


@pytest.mark.parametrize("mix,pseudo", [(0.1, 0.5), (0.0, 0.0), (1.0, 0.0)])
def test_learned_tables_match_restatement(mix, pseudo):
    c = mcmcmc.MCMCMCConfig(N1=2, N2=3)
    c.order_min, c.order_max, c.envelope_order, c.normalization_order = 1, 3, 4, 1
    c.base_seed = 2026101401
    c.histogram_learning_steps = 40
    c.histogram_tau_bins = 5
    c.histogram_uniform_mix, c.histogram_pseudo_count = mix, pseudo
    h, total = mcmcmc.learn_histogram(c, synthetic_fn(), None)
    T, X, R = py_learn(c, 5, 2, 3, mix, pseudo)
    np.testing.assert_allclose(h.tau, T, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(h.space, X, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(h.replica, R, rtol=1e-12, atol=1e-15)
    for M in (h.tau, h.space, h.replica):
        np.testing.assert_allclose(M.sum(axis=1), 1.0, rtol=1e-14)
    assert total > 0.0
    if mix == 1.0:
        assert np.allclose(h.tau, 1.0 / 5)


def test_learning_prefers_large_tau_for_growing_integrand():
    # integrand grows with every vertex's tau: children concentrate at large tau
    c = mcmcmc.MCMCMCConfig(N1=1, N2=1)
    c.order_min, c.order_max, c.envelope_order, c.normalization_order = 2, 4, 6, 2
    c.histogram_learning_steps, c.histogram_tau_bins, c.histogram_uniform_mix = 400, 4, 0.0
    h, _ = mcmcmc.learn_histogram(c, synthetic_fn(), None)
    col = h.tau.mean(axis=0)
    assert np.all(np.diff(col) > 0), col


def hist_config(d, **over):
    c = mcmcmc.MCMCMCConfig(N1=2, N2=2, beta=2.5)
    c.order_min, c.order_max, c.envelope_order, c.normalization_order = 1, 3, 4, 1
    for o in (1, 2, 3):
        c.lambdas[o] = 1.0
    c.lambda_warmup_steps, c.lambda_adapt_interval, c.lambda_adapt_rate = 200, 60, 0.2
    c.production_steps, c.chain_count, c.base_seed = 300, 2, 2026101401
    c.seed_mode = 1
    c.histogram_learning_steps, c.histogram_tau_bins = 64, 4
    c.checkpoint_enabled, c.checkpoint_interval_steps, c.resume_mode = 1, 100, 0
    c.checkpoint_dir = str(d).encode()
    for k, v in over.items():
        setattr(c, k, v.encode() if k in ("checkpoint_dir", "histogram_output_prefix") else v)
    return c


def oracle_eval():
    orc = helpers.oracle_of("ref2x2")
    return ctypes.cast(O.lib().orc_eval_batch_host, ctypes.c_void_p), orc


def chk_bytes(d):
    return {n: open(os.path.join(d, n), "rb").read() for n in sorted(os.listdir(d)) if n.endswith(".chk")}


def test_histogram_mode_resume_and_label(tmp_path):
    fn, orc = oracle_eval()
    full, cut = tmp_path / "full", tmp_path / "cut"
    r_full, _, t_full = mcmcmc.run(hist_config(full), fn, orc._h, trace_steps=500)
    mcmcmc.run(hist_config(cut, stop_after_steps=333), fn, orc._h)
    mcmcmc.run(hist_config(cut, resume_mode=2), fn, orc._h)
    a, b = chk_bytes(full), chk_bytes(cut)
    assert a == b and len(a) == 2
    assert all(b"seed_mode histogram\n" in x for x in a.values())
    # the learned seed changes the chain relative to the uniform one
    _, _, t_uni = mcmcmc.run(hist_config(tmp_path / "u", seed_mode=0, checkpoint_enabled=0), fn, orc._h,
                             trace_steps=500)
    assert not np.array_equal(t_full, t_uni)
    # run == learn + run_seeded
    h, _ = mcmcmc.learn_histogram(hist_config(tmp_path), fn, orc._h)
    _, _, t_seeded = mcmcmc.run_seeded(hist_config(tmp_path / "s", checkpoint_enabled=0), h, fn, orc._h,
                                       trace_steps=500)
    assert np.array_equal(t_full, t_seeded)


def test_host_threads_do_not_change_results(tmp_path):
    fn, orc = oracle_eval()
    outs = []
    for th in (1, 8):
        c = hist_config(tmp_path / f"t{th}", chain_count=8, host_threads=th, checkpoint_enabled=0)
        res, ratios, tr = mcmcmc.run(c, fn, orc._h, trace_steps=500)
        outs.append((tr, [list(r.order_signal_mean) for r in res], ratios))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]
    assert np.array_equal(outs[0][2], outs[1][2], equal_nan=True)


def test_learning_only_and_export(tmp_path):
    fn, orc = oracle_eval()
    prefix = tmp_path / "out" / "seed"
    c = hist_config(tmp_path, histogram_learning_only=1, histogram_output_prefix=str(prefix),
                    histogram_uniform_mix=0.25, histogram_pseudo_count=0.5, checkpoint_enabled=0)
    res, _, _ = mcmcmc.run(c, fn, orc._h)
    assert res[0].production_proposed_steps == 0  # no chain ran
    tau = np.loadtxt(str(prefix) + "_tau.csv", delimiter=",")
    space = np.loadtxt(str(prefix) + "_space.csv", delimiter=",")
    rep = np.loadtxt(str(prefix) + "_replica.csv", delimiter=",")
    assert tau.shape == (4, 4) and space.shape == (4, 4) and rep.shape == (2, 2)
    h, _ = mcmcmc.learn_histogram(c, fn, orc._h)
    assert np.array_equal(tau, h.tau) and np.array_equal(space, h.space)  # 17 significant digits round-trip
    meta = open(str(prefix) + "_metadata.txt").read()
    assert meta == "N1=2\nN2=2\nBETA=2.5\nTAU_BINS=4\nUNIFORM_MIX=0.25\nPSEUDO_COUNT=0.5\n"


@pytest.mark.parametrize("field,value,msg", [
    ("histogram_learning_steps", 0, "MCMCMC_SEED_HISTOGRAM_LEARNING_STEPS must be positive for histogram seed"),
    ("histogram_tau_bins", 0, "MCMCMC_SEED_HISTOGRAM_TAU_BINS must be positive"),
    ("histogram_uniform_mix", 1.5, r"MCMCMC_SEED_HISTOGRAM_UNIFORM_MIX must lie in \[0, 1\]"),
    ("histogram_pseudo_count", float("inf"), "MCMCMC_SEED_HISTOGRAM_PSEUDO_COUNT must be nonnegative and finite"),
    ("histogram_pseudo_count", -1.0, "MCMCMC_SEED_HISTOGRAM_PSEUDO_COUNT must be nonnegative and finite"),
])
def test_histogram_validation_messages(tmp_path, field, value, msg):
    fn, orc = oracle_eval()
    with pytest.raises(RuntimeError, match=msg):
        mcmcmc.run(hist_config(tmp_path, **{field: value}), fn, orc._h)


def test_defaults_match_reference():
    d = mcmcmc.defaults()
    assert (d.order_min, d.order_max, d.envelope_order, d.normalization_order) == (2, 4, 5, 2)
    assert (d.lambda_adapt_interval, d.lambda_adapt_rate, d.production_steps, d.chain_count) == (1000, 0.25, 100000, 1)
    assert (d.histogram_tau_bins, d.histogram_uniform_mix, d.histogram_pseudo_count) == (16, 0.05, 0.0)
    assert d.checkpoint_interval_steps == 10_000_000 and d.checkpoint_dir == b"mcmcmc_checkpoints"
    py = mcmcmc.MCMCMCConfig()
    assert (py.histogram_tau_bins, py.histogram_uniform_mix, py.N1, py.N2) == (16, 0.05, 1, 1)


def test_params_histogram_keys(tmp_path):
    p = tmp_path / "params.txt"
    p.write_text("N1=3\nN2=2\nBETA=4\nMCMCMC_SEED_MODE=histogram\nMCMCMC_SEED_HISTOGRAM_LEARNING_STEPS=77\n"
                 "MCMCMC_SEED_HISTOGRAM_TAU_BINS=8\nMCMCMC_SEED_HISTOGRAM_UNIFORM_MIX=0.3\n"
                 "MCMCMC_SEED_HISTOGRAM_PSEUDO_COUNT=1.5\nMCMCMC_SEED_HISTOGRAM_LEARNING_ONLY=true\n"
                 "MCMCMC_SEED_HISTOGRAM_OUTPUT_PREFIX=hist/run\nMCMCMC_COMPARE_UNIFORM_SEED=true\n")
    params = mcmcmc.read_params(str(p))
    c = mcmcmc.config_from_params(params)
    assert (c.seed_mode, c.N1, c.N2, c.beta) == (1, 3, 2, 4.0)
    assert (c.histogram_learning_steps, c.histogram_tau_bins, c.histogram_uniform_mix,
            c.histogram_pseudo_count, c.histogram_learning_only) == (77, 8, 0.3, 1.5, 1)
    assert c.histogram_output_prefix == b"hist/run" and mcmcmc.compare_uniform_seed(params)


def test_conditional_callback_matches_host_recursion(tmp_path):
    # a conditional-generation callback (here the host recursion through the
    # C-ABI, as the GPU backend provides it) gives the same chains as the
    # driver's own recursion
    fn, orc = oracle_eval()
    calls = []

    @mcmcmc.COND_FN
    def cond(user, seedp, verts, orders, count, out):
        h = ctypes.cast(seedp, ctypes.POINTER(mcmcmc.HistogramSeed)).contents
        vb = 0
        ob = 0
        for i in range(count):
            E = orders[i]
            v = np.ctypeslib.as_array(ctypes.cast(verts, ctypes.POINTER(ctypes.c_double)),
                                      shape=((vb + E) * 4,))[vb * 4:].reshape(E, 4)
            p = mcmcmc.conditional_probabilities(h, v)
            for k in range(1 << E):
                out[ob + k] = p[k]
            vb += E
            ob += 1 << E
        calls.append(count)
        return 0

    c = hist_config(tmp_path / "a", checkpoint_enabled=0, production_steps=120, lambda_warmup_steps=60)
    h, _ = mcmcmc.learn_histogram(c, fn, orc._h)
    _, _, t_host = mcmcmc.run_backend(c, fn, orc._h, seed=h, envelopes=[4, 3], trace_steps=200)
    _, _, t_cb = mcmcmc.run_backend(c, fn, orc._h, cond_fn=cond, seed=h, envelopes=[4, 3], trace_steps=200)
    assert np.array_equal(t_host, t_cb) and len(calls) == 180
    # no callback for the uniform seed
    calls.clear()
    _, _, t_u = mcmcmc.run_backend(c, fn, orc._h, cond_fn=cond, trace_steps=200)
    _, _, t_u2 = mcmcmc.run_seeded(c, None, fn, orc._h, trace_steps=200)
    assert np.array_equal(t_u, t_u2) and not calls

"""The lockstep driver's chains against a step-by-step restatement in Python of
the reference's sampler (include/cdet/mcmcmc_sampler.h), on the oracle's
integrand values (2x2 tables of the small reference set):

* initial state (make_initial_mcmcmc_chain_state, :697-709): mt19937_64 seeded
  with derive_mcmcmc_seed(base, chain tag, chain), E vertices of four uniform
  draws each (sample_uniform_unit_vertex, :186-189);
* mask plan (make_mcmcmc_mask_plan, :411-431): masks 1 .. 2^E - 1 of orders
  order_min .. order_max in increasing order, 1 / C(E, order) per order;
* one step (run_mcmcmc_step, :524-582): fluxes, selection weights and their
  running sum, the weighted choice (sample_weighted_candidate_mask,
  :225-248), the order signals, and the next vertices
  (create_next_vertices_from_seed_into, :380-399: kept vertices in index
  order, then per regenerated vertex a uniformly drawn parent and, for the
  uniform seed, a uniform child);
* warmup and production (:1037-1083): lambda adaptation (adapt_mcmcmc_lambdas,
  :584-606) every adapt interval and once more for an incomplete last block,
  the production order signals in long double running statistics (:31-78).

The draws follow the C++ standard library the driver is built with:
uniform_real_distribution<double>(a, b) as (double(g) / 2^64) (b - a) + a
(rng.MT19937_64.uniform), uniform_int_distribution over [0, k) by Lemire's
nearly divisionless method (D. Lemire, "Fast Random Integer Generation in an
Interval", ACM TOMACS 29(1), 2019) on 64-bit draws.  A plausible slip in the
driver -- a dropped lambda, a signal summed in another order, a wrong parent
draw, an adaptation off by one block -- changes the trace or the statistics.
"""
import math

import numpy as np

import helpers
import mcmcmc
import rng as R
from oracle import oracle as O

CHAIN_TAG = 0x4D434D434D435348  # include/cdet/mcmcmc_sampler.h:1103


def uniform_int(g, k):
    """Uniform integer in [0, k) from 64-bit draws (Lemire's method)."""
    m = 1 << 64
    prod = int(g.next_u64()[0]) * k
    if prod % m < k:
        threshold = (m - k) % k
        while prod % m < threshold:
            prod = int(g.next_u64()[0]) * k
    return prod >> 64


def uniform(g, a, b):
    return float(g.uniform(a, b)[0])


def uniform_vertex(g):
    return [uniform(g, 0.0, 1.0) for _ in range(4)]  # tau, x, y, replica


class Stats:
    """RunningStatsLongDouble (include/cdet/mcmcmc_sampler.h:31-78)."""

    def __init__(self):
        self.count = 0
        self.mean = np.longdouble(0)
        self.m2 = np.longdouble(0)

    def add(self, v):
        v = np.longdouble(v)
        self.count += 1
        d = v - self.mean
        self.mean = self.mean + d / np.longdouble(self.count)
        self.m2 = self.m2 + d * (v - self.mean)


def restated_chain(cfg, k, values_of):
    E, omin, omax = cfg["E"], cfg["order_min"], cfg["order_max"]
    g = R.MT19937_64([R.derive_mcmcmc_seed(cfg["base_seed"], CHAIN_TAG, k)])
    verts = [uniform_vertex(g) for _ in range(E)]
    lambdas = list(cfg["lambdas"])
    masks = [m for m in range(1, 1 << E) if omin <= bin(m).count("1") <= omax]
    orders = [bin(m).count("1") for m in masks]
    inv_binom = [0.0] * (E + 1)
    for o in range(omin, omax + 1):
        inv_binom[o] = 1.0 / float(math.comb(E, o))
    block = [0] * (E + 1)
    trace, warm_chosen, prod_chosen = [], [0] * (E + 1), [0] * (E + 1)
    signal = [Stats() for _ in range(E + 1)]

    def adapt():
        total = sum(block[o] for o in range(omin, omax + 1))
        if total == 0:
            return
        oc = omax - omin + 1
        target = 1.0 / float(oc)
        for o in range(omin, omax + 1):
            actual = (float(block[o]) + 1.0) / (float(total) + float(oc))
            lambdas[o] *= math.exp(cfg["rate"] * math.log(target / actual))

    def step():
        nonlocal verts
        vals = values_of(np.array(verts))
        fluxes, weights, norm = [], [], 0.0
        for m, o in zip(masks, orders):
            flux = lambdas[o] * float(vals[m]) * 1.0
            w = abs(flux) * inv_binom[o]
            fluxes.append(flux)
            weights.append(w)
            norm += w
        assert norm > 0.0 and math.isfinite(norm)
        threshold = uniform(g, 0.0, norm)
        cumulative, last_positive, chosen = 0.0, 0, None
        for m, w in zip(masks, weights):
            if not w > 0.0:
                continue
            if math.isfinite(w):
                last_positive = m
            cumulative += w
            if cumulative > threshold:
                chosen = m
                break
        if chosen is None:
            chosen = last_positive
        osig = [0.0] * (E + 1)
        for f, o in zip(fluxes, orders):
            osig[o] += (f / norm) * inv_binom[o]
        nxt = [verts[i] for i in range(E) if chosen >> i & 1]
        while len(nxt) < E:
            uniform_int(g, len(nxt))  # parent; the uniform seed's child ignores it
            nxt.append(uniform_vertex(g))
        verts = nxt
        trace.append(chosen)
        return bin(chosen).count("1"), osig

    for s in range(cfg["warmup"]):
        o, _ = step()
        warm_chosen[o] += 1
        block[o] += 1
        if (s + 1) % cfg["interval"] == 0:
            adapt()
            block = [0] * (E + 1)
    if cfg["warmup"] % cfg["interval"] != 0 and any(block):
        adapt()
    for _ in range(cfg["production"]):
        o, osig = step()
        prod_chosen[o] += 1
        for q in range(omin, omax + 1):
            signal[q].add(osig[q] / lambdas[q])
    return trace, warm_chosen, prod_chosen, [float(s.mean) for s in signal]


def test_driver_steps_match_restated_sampler():
    orc = helpers.oracle_of("ref2x2")
    fn = O.lib().orc_eval_batch_host
    cfg = dict(E=5, order_min=1, order_max=4, base_seed=2026101401, warmup=70, interval=30, rate=0.2,
               production=90, lambdas=[0.0, 1.0, 0.8, 1.3, 0.6, 0.0])
    c = mcmcmc.MCMCMCConfig()
    c.order_min, c.order_max, c.envelope_order, c.normalization_order = cfg["order_min"], cfg["order_max"], cfg["E"], 1
    for o, lam in enumerate(cfg["lambdas"]):
        c.lambdas[o] = lam
    c.lambda_warmup_steps, c.lambda_adapt_interval, c.lambda_adapt_rate = cfg["warmup"], cfg["interval"], cfg["rate"]
    c.production_steps, c.chain_count, c.base_seed = cfg["production"], 3, cfg["base_seed"]
    c.seed_mode, c.checkpoint_enabled, c.resume_mode = 0, 0, 0
    steps = cfg["warmup"] + cfg["production"]
    res, _, trace = mcmcmc.run(c, fn, orc._h, trace_steps=steps)
    for k in range(3):
        t, wc, pc, means = restated_chain(cfg, k, orc.all_subsets)
        assert [int(x) for x in trace[k]] == t, k
        assert list(res[k].warmup_chosen_counts[:cfg["E"] + 1]) == wc, k
        assert list(res[k].production_chosen_counts[:cfg["E"] + 1]) == pc, k
        for o in range(cfg["order_min"], cfg["order_max"] + 1):
            assert res[k].order_signal_mean[o] == means[o], (k, o)

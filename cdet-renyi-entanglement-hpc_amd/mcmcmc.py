"""Host mirror of the reference's MCMCMC driver configuration and entry point.

Reads the reference's params.txt keys (src/main.cpp:218-576) for the sampler
and the integrand, and runs the lockstep driver of libcdet_mcmcmc.so
(include/cdet_mcmcmc.h) over any batch-evaluation callback: the GPU library
(cdet_eval_batch_host) in production, the oracle in tests.
"""
import ctypes
import os
from fractions import Fraction

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcdet_mcmcmc.so")
MAXE = 48
_lib = None

EVAL_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                           ctypes.c_int64, ctypes.POINTER(ctypes.c_double))


class MCMCMCConfig(ctypes.Structure):
    """cdet_mcmcmc_config (include/cdet_mcmcmc.h)."""

    _fields_ = [
        ("order_min", ctypes.c_int32), ("order_max", ctypes.c_int32),
        ("envelope_order", ctypes.c_int32), ("normalization_order", ctypes.c_int32),
        ("lambdas", ctypes.c_double * (MAXE + 1)),
        ("lambda_warmup_steps", ctypes.c_int32), ("lambda_adapt_interval", ctypes.c_int32),
        ("lambda_adapt_rate", ctypes.c_double),
        ("production_steps", ctypes.c_int32), ("chain_count", ctypes.c_int32),
        ("base_seed", ctypes.c_uint64), ("seed_mode", ctypes.c_int32),
        ("checkpoint_enabled", ctypes.c_int32), ("checkpoint_interval_steps", ctypes.c_uint64),
        ("resume_mode", ctypes.c_int32), ("checkpoint_dir", ctypes.c_char * 1024),
        ("stop_after_steps", ctypes.c_uint64),
        ("N1", ctypes.c_int32), ("N2", ctypes.c_int32), ("beta", ctypes.c_double),
        ("histogram_learning_steps", ctypes.c_int32), ("histogram_tau_bins", ctypes.c_int32),
        ("histogram_uniform_mix", ctypes.c_double), ("histogram_pseudo_count", ctypes.c_double),
        ("histogram_learning_only", ctypes.c_int32), ("histogram_output_prefix", ctypes.c_char * 1024),
        ("host_threads", ctypes.c_int32),
    ]

    def __init__(self, **kw):
        super().__init__()
        # The histogram settings are validated in every seed mode; start from
        # the reference's defaults (include/cdet/mcmcmc_sampler.h:92-95).
        self.N1 = self.N2 = 1
        self.beta = 1.0
        self.histogram_tau_bins = 16
        self.histogram_uniform_mix = 0.05
        for k, v in kw.items():
            setattr(self, k, v)


class HistogramSeed(ctypes.Structure):
    """cdet_histogram_seed (include/cdet_mcmcmc.h); the numpy tables are kept
    alive by the instance."""

    _fields_ = [
        ("tau_bins", ctypes.c_int32), ("N1", ctypes.c_int32), ("N2", ctypes.c_int32),
        ("uniform_mix", ctypes.c_double), ("pseudo_count", ctypes.c_double),
        ("tau_prob", ctypes.POINTER(ctypes.c_double)), ("space_prob", ctypes.POINTER(ctypes.c_double)),
        ("replica_prob", ctypes.c_double * 4),
    ]

    @classmethod
    def make(cls, tau_bins, N1, N2, tau=None, space=None, replica=None, uniform_mix=0.0, pseudo_count=0.0):
        h = cls()
        h.tau_bins, h.N1, h.N2 = tau_bins, N1, N2
        h.uniform_mix, h.pseudo_count = uniform_mix, pseudo_count
        S = N1 * N2
        h._tau = np.ascontiguousarray(np.zeros((tau_bins, tau_bins)) if tau is None else tau, dtype=np.float64)
        h._space = np.ascontiguousarray(np.zeros((S, S)) if space is None else space, dtype=np.float64)
        assert h._tau.shape == (tau_bins, tau_bins) and h._space.shape == (S, S)
        h.tau_prob = h._tau.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        h.space_prob = h._space.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        if replica is not None:
            for i, x in enumerate(np.asarray(replica, dtype=np.float64).ravel()):
                h.replica_prob[i] = x
        return h

    @property
    def tau(self):
        return self._tau

    @property
    def space(self):
        return self._space

    @property
    def replica(self):
        return np.array(self.replica_prob[:]).reshape(2, 2)


COND_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                           ctypes.POINTER(ctypes.c_int32), ctypes.c_int64, ctypes.POINTER(ctypes.c_double))


class Backend(ctypes.Structure):
    """cdet_mcmcmc_backend (include/cdet_mcmcmc.h)."""

    _fields_ = [("eval", ctypes.c_void_p), ("eval_user", ctypes.c_void_p),
                ("cond", ctypes.c_void_p), ("cond_user", ctypes.c_void_p),
                ("submit", ctypes.c_void_p), ("wait", ctypes.c_void_p),
                ("host_alloc", ctypes.c_void_p), ("host_free", ctypes.c_void_p)]


# cdet_batch_submit_fn / cdet_batch_wait_fn, for callbacks written in Python
SUBMIT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                             ctypes.POINTER(ctypes.c_int32), ctypes.c_int64)
WAIT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double))


class ChainResult(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64), ("warmup_proposed_steps", ctypes.c_uint64),
        ("production_proposed_steps", ctypes.c_uint64),
        ("warmup_chosen_counts", ctypes.c_uint64 * (MAXE + 1)),
        ("production_chosen_counts", ctypes.c_uint64 * (MAXE + 1)),
        ("final_lambdas", ctypes.c_double * (MAXE + 1)),
        ("order_signal_mean", ctypes.c_double * (MAXE + 1)),
        ("order_signal_stderr", ctypes.c_double * (MAXE + 1)),
        ("order_signal_count", ctypes.c_uint64 * (MAXE + 1)),
        ("retained_fraction_mean", ctypes.c_double),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcdet_mcmcmc.so is not built; run cdet-renyi-entanglement-hpc_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        L.cdet_mcmcmc_run.restype = ctypes.c_int
        L.cdet_mcmcmc_run.argtypes = [ctypes.POINTER(MCMCMCConfig), ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.POINTER(ChainResult), ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_uint64), ctypes.c_int64]
        L.cdet_mcmcmc_last_error.restype = ctypes.c_char_p
        L.cdet_mcmcmc_config_defaults.restype = None
        L.cdet_mcmcmc_config_defaults.argtypes = [ctypes.POINTER(MCMCMCConfig)]
        L.cdet_mcmcmc_run_seeded.restype = ctypes.c_int
        L.cdet_mcmcmc_run_seeded.argtypes = [ctypes.POINTER(MCMCMCConfig), ctypes.POINTER(HistogramSeed),
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ChainResult),
                                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64),
                                             ctypes.c_int64]
        L.cdet_mcmcmc_run_envelopes.restype = ctypes.c_int
        L.cdet_mcmcmc_run_envelopes.argtypes = [ctypes.POINTER(MCMCMCConfig), ctypes.POINTER(HistogramSeed),
                                                ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.POINTER(ChainResult), ctypes.POINTER(ctypes.c_double),
                                                ctypes.POINTER(ctypes.c_uint64), ctypes.c_int64]
        L.cdet_mcmcmc_run_backend.restype = ctypes.c_int
        L.cdet_mcmcmc_run_backend.argtypes = [ctypes.POINTER(MCMCMCConfig), ctypes.POINTER(HistogramSeed),
                                              ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(Backend),
                                              ctypes.POINTER(ChainResult), ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_uint64), ctypes.c_int64]
        L.cdet_histogram_learn.restype = ctypes.c_int
        L.cdet_histogram_learn.argtypes = [ctypes.POINTER(MCMCMCConfig), ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.POINTER(HistogramSeed), ctypes.POINTER(ctypes.c_double)]
        L.cdet_histogram_export.restype = ctypes.c_int
        L.cdet_histogram_export.argtypes = [ctypes.POINTER(HistogramSeed), ctypes.c_double, ctypes.c_char_p]
        L.cdet_seed_conditional_probabilities.restype = ctypes.c_int
        L.cdet_seed_conditional_probabilities.argtypes = [ctypes.POINTER(HistogramSeed), ctypes.c_void_p,
                                                          ctypes.c_int, ctypes.c_void_p]
        L.cdet_seed_sample_children.restype = ctypes.c_int
        L.cdet_seed_sample_children.argtypes = [ctypes.POINTER(HistogramSeed), ctypes.c_void_p, ctypes.c_int64,
                                                ctypes.c_uint64, ctypes.c_void_p]
        L.cdet_seed_density.restype = ctypes.c_int
        L.cdet_seed_density.argtypes = [ctypes.POINTER(HistogramSeed), ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p]
        _lib = L
    return _lib


# ---- params.txt (src/main.cpp:218-261: KEY=value[,value...], '#' comments) ----
def read_params(path):
    params = {}
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line or "=" not in line:
                continue
            k, v = line.split("=", 1)
            params[k.strip()] = [x.strip() for x in v.split(",")]
    return params


def _one(params, key, default=None):
    if key in params:
        return params[key][0]
    if default is None:
        raise KeyError(f"missing required parameter {key}")
    return default


def _bool(text):
    return text.strip().lower() in ("1", "true", "yes", "on")


def config_from_params(params, **over):
    """MCMCMCConfig from params.txt keys and defaults (src/main.cpp:461-535)."""
    omin = int(_one(params, "MCMCMC_ORDER_MIN", "2"))
    omax = int(_one(params, "MCMCMC_ORDER_MAX", "4"))
    env = int(_one(params, "MCMCMC_ENVELOPE_ORDER", str(omax + 1)))
    norm = int(_one(params, "MCMCMC_NORMALIZATION_ORDER", str(omin)))
    lam = [float(x) for x in params.get("MCMCMC_LAMBDAS", ["1"] * (omax - omin + 1))]
    if len(lam) == 1:
        lam = lam * (omax - omin + 1)
    c = MCMCMCConfig()
    c.order_min, c.order_max, c.envelope_order, c.normalization_order = omin, omax, env, norm
    for o in range(omin, omax + 1):
        c.lambdas[o] = lam[o - omin]
    c.lambda_warmup_steps = int(_one(params, "MCMCMC_LAMBDA_WARMUP_STEPS", "10000"))
    c.lambda_adapt_interval = int(_one(params, "MCMCMC_LAMBDA_ADAPT_INTERVAL", "1000"))
    c.lambda_adapt_rate = float(_one(params, "MCMCMC_LAMBDA_ADAPT_RATE", "0.25"))
    c.production_steps = int(_one(params, "MCMCMC_PRODUCTION_STEPS", "100000"))
    c.chain_count = int(_one(params, "MCMCMC_CHAIN_COUNT", "8"))
    c.base_seed = int(_one(params, "MCMCMC_BASE_SEED", "998875432210"))
    c.seed_mode = {"uniform": 0, "histogram": 1}[_one(params, "MCMCMC_SEED_MODE", "uniform").lower()]
    c.checkpoint_enabled = int(_bool(_one(params, "MCMCMC_CHECKPOINT_ENABLED", "false")))
    c.checkpoint_interval_steps = int(_one(params, "MCMCMC_CHECKPOINT_INTERVAL_STEPS", "10000000"))
    rm = _one(params, "MCMCMC_RESUME", "off").lower()
    c.resume_mode = {"off": 0, "auto": 1, "require": 2, "required": 2}[rm]
    c.checkpoint_dir = _one(params, "MCMCMC_CHECKPOINT_DIR", "mcmcmc_checkpoints").encode()
    # Histogram seed keys (src/main.cpp:495-509).
    c.N1 = int(_one(params, "N1", "1"))
    c.N2 = int(_one(params, "N2", "1"))
    c.beta = float(_one(params, "BETA", "1"))
    c.histogram_learning_steps = int(_one(params, "MCMCMC_SEED_HISTOGRAM_LEARNING_STEPS", "0"))
    c.histogram_tau_bins = int(_one(params, "MCMCMC_SEED_HISTOGRAM_TAU_BINS", "16"))
    c.histogram_uniform_mix = float(_one(params, "MCMCMC_SEED_HISTOGRAM_UNIFORM_MIX", "0.05"))
    c.histogram_pseudo_count = float(_one(params, "MCMCMC_SEED_HISTOGRAM_PSEUDO_COUNT", "0"))
    c.histogram_learning_only = int(_bool(_one(params, "MCMCMC_SEED_HISTOGRAM_LEARNING_ONLY", "false")))
    c.histogram_output_prefix = _one(params, "MCMCMC_SEED_HISTOGRAM_OUTPUT_PREFIX", "").encode()
    for k, v in over.items():
        setattr(c, k, v.encode() if k in ("checkpoint_dir", "histogram_output_prefix") else v)
    return c


def _normalize_name(name):
    """normalize_real_type_name (include/cdet/precision_config.h:28-36)."""
    return name.lower().replace("-", "_")


def table_workflow_from_params(params, environ=None):
    """LOAD_DATA_WORKFLOW, MCMCMC_TABLE_BACKEND and LOAD_DATA_CHUNK_DIR with
    their environment overrides and aliases (src/main.cpp:322-345, :428-446).
    Returns (workflow, backend, chunk_dir); raises RuntimeError with the
    reference's message for an unknown backend."""
    env = os.environ if environ is None else environ
    workflow = _normalize_name(_one(params, "LOAD_DATA_WORKFLOW", "resident"))
    if env.get("CDET_LOAD_DATA_WORKFLOW"):
        workflow = _normalize_name(env["CDET_LOAD_DATA_WORKFLOW"])
    workflow = {"build_resident": "resident", "load_prebuilt": "load_chunks",
                "load_prebuilt_coeffs": "load_chunks", "sample": "sample_nodes",
                "precompute": "precompute_calculators", "calculator_cache": "precompute_calculators",
                "transform": "transform_sitepairs", "diagonal_observable": "diagonal_reduce",
                "diagonal_lines": "diagonal_reduce"}.get(workflow, workflow)
    backend = _normalize_name(_one(params, "MCMCMC_TABLE_BACKEND", "resident"))
    if env.get("CDET_MCMCMC_TABLE_BACKEND"):
        backend = _normalize_name(env["CDET_MCMCMC_TABLE_BACKEND"])
    if backend in ("lru", "on_demand", "disk", "disk_cache"):
        backend = "disk_lru"
    if backend not in ("resident", "disk_lru"):
        raise RuntimeError("MCMCMC_TABLE_BACKEND must be resident or disk_lru")
    chunk_dir = _one(params, "LOAD_DATA_CHUNK_DIR", "chunked_green_tables")
    if env.get("CDET_LOAD_DATA_CHUNK_DIR"):
        chunk_dir = env["CDET_LOAD_DATA_CHUNK_DIR"]
    return workflow, backend, chunk_dir


def load_data_from_params(ctx, params, environ=None):
    """load_data() (include/cdet/load_data.h:119-178) onto a GPU context.

    LOAD_DATA_WORKFLOW=load_chunks streams the chunk directory's coefficient
    files into HBM.  Both table backends are accepted: `resident` as the
    reference means it, and `disk_lru` (coefficient chunks kept on disk
    behind an LRU cache, include/cdet/replica_chebyshev_tables.h:56-152) maps
    to the same full residency in HBM, since the largest tables of this path
    (10x10 lattice, 192 coefficients: 4 x 23.6 GB) fit one GPU's 288 GB; an
    MCMCMC_LRU_CACHE_GB budget is accepted and not needed.
    LOAD_DATA_WORKFLOW=resident builds the tables on the GPU.  Other workflows
    and disk_lru without load_chunks raise the reference's errors.
    Returns {"workflow", "backend", "residency"}."""
    workflow, backend, chunk_dir = table_workflow_from_params(params, environ)
    if workflow == "load_chunks":
        ctx.load_chunks(chunk_dir)
        return {"workflow": workflow, "backend": backend, "residency": "hbm"}
    if backend == "disk_lru":
        raise RuntimeError("MCMCMC_TABLE_BACKEND=disk_lru requires LOAD_DATA_WORKFLOW=load_chunks "
                           "so coefficient chunks can be read on demand")
    if workflow != "resident":
        raise RuntimeError("load_data() can only run with LOAD_DATA_WORKFLOW=resident or load_chunks; "
                           "standalone chunk stages must exit before MCMCMC starts")
    mask = _normalize_name(_one(params, "SUBREGION_MASK"))
    radius = int(_one(params, "SUBREGION_SQUARE_RADIUS", "0"))
    ctx.build_tables(pbc=_bool(_one(params, "PBC")), mask=mask,
                     mask_param=radius if mask == "origin_square" else 0)
    return {"workflow": workflow, "backend": backend, "residency": "hbm"}


def compare_uniform_seed(params):
    """MCMCMC_COMPARE_UNIFORM_SEED (src/main.cpp:508-509)."""
    return _bool(_one(params, "MCMCMC_COMPARE_UNIFORM_SEED", "false"))


def integrand_scalars_from_params(params):
    """Derived integrand scalars from the decimal parameter text, correctly
    rounded (the reference forms them in double-double from 100-digit values,
    include/cdet/integrand_replica_randomly_sampled.h:98-119)."""
    from decimal import Decimal, getcontext
    getcontext().prec = 60
    N1 = int(_one(params, "N1"))
    N2 = int(_one(params, "N2"))
    U = Fraction(_one(params, "U"))
    beta = Fraction(_one(params, "BETA"))
    a_up = Fraction(_one(params, "ALPHA_UP"))
    a_dn = Fraction(_one(params, "ALPHA_DOWN"))
    prod = U * N1 * N2 * beta
    scale = float(Fraction((Decimal(prod.numerator) / Decimal(prod.denominator)).sqrt()))
    return scale, float(a_dn / U), float(a_up / U)


def cross_chain_stats(results, order_min, order_max):
    """Cross-chain statistics of each order's signal: every chain's mean is one
    independent CLT sample (include/cdet/mcmcmc_sampler.h:1117-1131, with the
    running mean and M2 of RunningStatsLongDouble, :31-78).  Returns
    {order: (mean, stderr, chains)}; stderr is NaN below two chains."""
    out = {}
    for order in range(order_min, order_max + 1):
        count, mean, m2 = 0, 0.0, 0.0
        for r in results:
            if r.order_signal_count[order] > 0:
                count += 1
                x = float(r.order_signal_mean[order])
                delta = x - mean
                mean += delta / count
                m2 += delta * (x - mean)
        se = float(np.sqrt(m2 / (count - 1) / count)) if count >= 2 else float("nan")
        out[order] = (mean, se, count)
    return out


def agree_within(stats_a, stats_b, sigmas=2.0):
    """Per order: |mean_a - mean_b| <= sigmas * sqrt(se_a^2 + se_b^2) (orders
    whose signal is identically zero in both count as agreeing).  Returns
    {order: (difference / combined stderr, agrees)}."""
    out = {}
    for order in stats_a:
        ma, sa, _ = stats_a[order]
        mb, sb, _ = stats_b[order]
        se = float(np.hypot(sa, sb))
        d = abs(ma - mb)
        z = d / se if se > 0 else (0.0 if d == 0 else float("inf"))
        out[order] = (z, bool(z <= sigmas))
    return out


def run(config, eval_fn, eval_user, trace_steps=0):
    """Run the lockstep driver; eval_fn is a ctypes function pointer (or an
    address) of cdet_batch_eval_fn type."""
    n = config.chain_count
    results = (ChainResult * n)()
    ratios = (ctypes.c_double * (MAXE + 1))()
    trace = np.zeros((n, max(trace_steps, 1)), dtype=np.uint64)
    fn = ctypes.cast(eval_fn, ctypes.c_void_p)
    rc = lib().cdet_mcmcmc_run(ctypes.byref(config), fn, eval_user, results, ratios,
                               trace.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)) if trace_steps else None,
                               trace_steps)
    if rc != 0:
        raise RuntimeError(lib().cdet_mcmcmc_last_error().decode())
    return list(results), np.array(ratios[:config.envelope_order + 1]), (trace if trace_steps else None)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().cdet_mcmcmc_last_error().decode())


def run_seeded(config, seed, eval_fn, eval_user, trace_steps=0):
    """Run the chains with a given seed kernel (None = uniform)."""
    n = config.chain_count
    results = (ChainResult * n)()
    ratios = (ctypes.c_double * (MAXE + 1))()
    trace = np.zeros((n, max(trace_steps, 1)), dtype=np.uint64)
    fn = ctypes.cast(eval_fn, ctypes.c_void_p)
    _check(lib().cdet_mcmcmc_run_seeded(
        ctypes.byref(config), ctypes.byref(seed) if seed is not None else None, fn, eval_user, results, ratios,
        trace.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)) if trace_steps else None, trace_steps))
    return list(results), np.array(ratios[:config.envelope_order + 1]), (trace if trace_steps else None)


def run_envelopes(config, envelopes, eval_fn, eval_user, seed=None, trace_steps=0):
    """Chains with their own envelope orders (cdet_mcmcmc_run_envelopes)."""
    n = config.chain_count
    env = np.ascontiguousarray(envelopes, dtype=np.int32)
    assert env.shape == (n,)
    results = (ChainResult * n)()
    ratios = (ctypes.c_double * (MAXE + 1))()
    trace = np.zeros((n, max(trace_steps, 1)), dtype=np.uint64)
    _check(lib().cdet_mcmcmc_run_envelopes(
        ctypes.byref(config), ctypes.byref(seed) if seed is not None else None,
        env.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.cast(eval_fn, ctypes.c_void_p), eval_user,
        results, ratios, trace.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)) if trace_steps else None,
        trace_steps))
    return list(results), np.array(ratios[:int(env.max()) + 1]), (trace if trace_steps else None)


def run_backend(config, eval_fn, eval_user, cond_fn=None, cond_user=None, seed=None, envelopes=None,
                trace_steps=0, submit_fn=None, wait_fn=None, host_alloc_fn=None, host_free_fn=None):
    """Chains through a backend (cdet_mcmcmc_run_backend): integrand callback
    and, optionally, the conditional-generation callback of the histogram
    seed, the asynchronous submit/wait pair (two groups of chains in flight)
    and the allocator of the value buffers (cdet_host_alloc / cdet_host_free
    for the GPU backend).  envelopes: per-chain envelope orders or None."""
    n = config.chain_count
    env = None if envelopes is None else np.ascontiguousarray(envelopes, dtype=np.int32)
    cast = lambda f: ctypes.cast(f, ctypes.c_void_p) if f is not None else None
    be = Backend(cast(eval_fn), eval_user, cast(cond_fn), cond_user, cast(submit_fn), cast(wait_fn),
                 cast(host_alloc_fn), cast(host_free_fn))
    results = (ChainResult * n)()
    ratios = (ctypes.c_double * (MAXE + 1))()
    trace = np.zeros((n, max(trace_steps, 1)), dtype=np.uint64)
    _check(lib().cdet_mcmcmc_run_backend(
        ctypes.byref(config), ctypes.byref(seed) if seed is not None else None,
        env.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if env is not None else None, ctypes.byref(be),
        results, ratios, trace.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)) if trace_steps else None,
        trace_steps))
    emax = config.envelope_order if env is None else int(env.max())
    return list(results), np.array(ratios[:emax + 1]), (trace if trace_steps else None)


def learn_histogram(config, eval_fn, eval_user):
    """Learn the factorized histogram seed; returns (HistogramSeed, total pair weight)."""
    h = HistogramSeed.make(config.histogram_tau_bins, config.N1, config.N2)
    tot = ctypes.c_double()
    _check(lib().cdet_histogram_learn(ctypes.byref(config), ctypes.cast(eval_fn, ctypes.c_void_p), eval_user,
                                      ctypes.byref(h), ctypes.byref(tot)))
    return h, tot.value


def export_histogram(seed, beta, prefix):
    _check(lib().cdet_histogram_export(ctypes.byref(seed), beta, prefix.encode()))


def conditional_probabilities(seed, vertices):
    v = np.ascontiguousarray(vertices, dtype=np.float64)
    E = v.shape[0]
    out = np.zeros(1 << E)
    _check(lib().cdet_seed_conditional_probabilities(ctypes.byref(seed) if seed is not None else None,
                                                     v.ctypes.data, E, out.ctypes.data))
    return out


def sample_children(seed, parents, rng_seed):
    p = np.ascontiguousarray(parents, dtype=np.float64).reshape(-1, 4)
    out = np.zeros_like(p)
    _check(lib().cdet_seed_sample_children(ctypes.byref(seed) if seed is not None else None, p.ctypes.data,
                                           p.shape[0], rng_seed, out.ctypes.data))
    return out


def seed_density(seed, child, parent):
    c = np.ascontiguousarray(child, dtype=np.float64)
    p = np.ascontiguousarray(parent, dtype=np.float64)
    out = np.zeros(1)
    _check(lib().cdet_seed_density(ctypes.byref(seed) if seed is not None else None, c.ctypes.data,
                                   p.ctypes.data, out.ctypes.data))
    return float(out[0])


def defaults():
    c = MCMCMCConfig()
    lib().cdet_mcmcmc_config_defaults(ctypes.byref(c))
    return c

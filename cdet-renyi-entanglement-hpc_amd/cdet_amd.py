"""Host-side mirror of the reference's integrand API over the gfx950 library.

The reference evaluates the CDet integrand in-process from C++ templates
(include/cdet/cdet_integrand_evaluator.h:542-590).  This module exposes the
same calls, with the same argument meaning and error behaviour, over the
C-ABI of libcdet_hip.so (include/cdet_hip.h):

  Context.evaluate_integrand(vertices)               -> float     (:542-560)
  Context.evaluate_all_subset_integrands(vertices)   -> ndarray   (:562-590)
  BatchPlan.run(d_vertices, d_out, d_status)                        (batched)

Errors the reference raises as std::runtime_error are raised here as
RuntimeError with the reference's message.  There is no CPU fallback: if the
HIP library is missing, loading fails loudly.
"""
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcdet_hip.so")

MODE_EXACT = 0
MODE_MFMA64 = 1
MODE_SPLIT32 = 2
MODES = {"exact": MODE_EXACT, "mfma64": MODE_MFMA64, "split32": MODE_SPLIT32}
SUM_ORDERS = {"tree": 0, "reference": 1}  # include/cdet_hip.h CDET_SUM_*

TWISTED_UP, TWISTED_DOWN, ORDINARY_UP, ORDINARY_DOWN = 0, 1, 2, 3
MASK_KINDS = {"empty": 0, "checkerboard": 1, "first_row": 2, "first_column": 3,
              "origin_square": 4, "left_columns": 5}

_lib = None


class CdetParams(ctypes.Structure):
    """cdet_params (include/cdet_hip.h)."""

    _fields_ = [
        ("N1", ctypes.c_int32), ("N2", ctypes.c_int32),
        ("cheb_nu", ctypes.c_int32), ("cheb_nv", ctypes.c_int32),
        ("beta", ctypes.c_double), ("U", ctypes.c_double), ("mu", ctypes.c_double),
        ("alpha_up", ctypes.c_double), ("alpha_down", ctypes.c_double),
        ("branch_epsilon", ctypes.c_double), ("tol_fpm", ctypes.c_double),
        ("fpm_global_shift", ctypes.c_double),
        ("ztransform_epsilon0", ctypes.c_double), ("ztransform_safety", ctypes.c_double),
        ("scale", ctypes.c_double), ("diag_up_shift", ctypes.c_double),
        ("diag_down_shift", ctypes.c_double),
    ]


MAX_ORDER = 14


class CdetTuning(ctypes.Structure):
    """cdet_tuning (include/cdet_hip.h): launch shapes; 0 keeps a default."""

    _fields_ = [
        ("contract_wg_per_cu", ctypes.c_int32), ("contract_tile", ctypes.c_int32),
        ("pair_min_per_cu", ctypes.c_int32), ("max_lanes", ctypes.c_int32),
        ("minors_threads", ctypes.c_int32 * (MAX_ORDER + 1)),
        ("group16_min_terms", ctypes.c_int32 * (MAX_ORDER + 1)),
        ("group64_min_terms", ctypes.c_int32 * (MAX_ORDER + 1)),
        ("fused_minors_max_order", ctypes.c_int32),
    ]

    SCALARS = ("contract_wg_per_cu", "contract_tile", "pair_min_per_cu", "max_lanes", "fused_minors_max_order")
    PER_ORDER = ("minors_threads", "group16_min_terms", "group64_min_terms")

    def to_dict(self):
        d = {k: int(getattr(self, k)) for k in self.SCALARS}
        for k in self.PER_ORDER:
            d[k] = [int(x) for x in getattr(self, k)]
        return d

    @classmethod
    def from_dict(cls, d):
        t = cls()
        for k in cls.SCALARS:
            setattr(t, k, int(d.get(k, 0)))
        for k in cls.PER_ORDER:
            vals = d.get(k, [])
            arr = getattr(t, k)
            if isinstance(vals, dict):  # {order: value}
                for n, v in vals.items():
                    arr[int(n)] = int(v)
            else:
                for n, v in enumerate(vals[:MAX_ORDER + 1]):
                    arr[n] = int(v)
        return t


class CdetCounters(ctypes.Structure):
    _fields_ = [("integrand_calls", ctypes.c_uint64), ("fpm_fallbacks", ctypes.c_uint64),
                ("green_pair_cache_hits", ctypes.c_uint64), ("green_pair_cache_misses", ctypes.c_uint64)]


class CdetLattice(ctypes.Structure):
    _fields_ = [("pbc", ctypes.c_int32), ("mask_kind", ctypes.c_int32),
                ("mask_param", ctypes.c_int32), ("node_chunk", ctypes.c_int32),
                ("ordinary_layout", ctypes.c_int32)]


ORDINARY_RESIDENT, ORDINARY_CHUNKED = 0, 1


class ChunkSpec(ctypes.Structure):
    """cdet_chunk_spec (include/cdet_chunks.h)."""

    _fields_ = [("N1", ctypes.c_int32), ("N2", ctypes.c_int32),
                ("cheb_nu", ctypes.c_int32), ("cheb_nv", ctypes.c_int32),
                ("node_chunk_size", ctypes.c_int64), ("sitepair_chunk_size", ctypes.c_int64),
                ("beta", ctypes.c_double), ("mu", ctypes.c_double),
                ("alpha_up", ctypes.c_double), ("alpha_down", ctypes.c_double),
                ("pbc", ctypes.c_int32), ("mask_kind", ctypes.c_int32),
                ("subregion_square_radius", ctypes.c_int32)]


def lattice(pbc=True, mask="first_row", mask_param=0, node_chunk=0, ordinary_layout=ORDINARY_RESIDENT):
    kind = MASK_KINDS[mask] if isinstance(mask, str) else int(mask)
    return CdetLattice(pbc=int(pbc), mask_kind=kind, mask_param=mask_param, node_chunk=node_chunk,
                       ordinary_layout=ordinary_layout)


def chunk_spec(params, node_chunk_size, sitepair_chunk_size, pbc=True, mask="first_row", radius=0):
    kind = MASK_KINDS[mask] if isinstance(mask, str) else int(mask)
    return ChunkSpec(N1=params.N1, N2=params.N2, cheb_nu=params.cheb_nu, cheb_nv=params.cheb_nv,
                     node_chunk_size=node_chunk_size, sitepair_chunk_size=sitepair_chunk_size,
                     beta=params.beta, mu=params.mu, alpha_up=params.alpha_up, alpha_down=params.alpha_down,
                     pbc=int(pbc), mask_kind=kind, subregion_square_radius=radius)


def chunks_write_plan(directory, spec):
    """Job plans and manifest of a chunk directory (host only)."""
    _check(lib().cdet_chunks_write_plan(str(directory).encode(), ctypes.byref(spec)))


def chunks_read_spec(directory):
    spec = ChunkSpec()
    _check(lib().cdet_chunks_read_spec(str(directory).encode(), ctypes.byref(spec)))
    return spec


# Exported symbols and their signatures, as declared in include/cdet_hip.h.
def _signatures():
    P = ctypes.POINTER
    i, i64, d, vp = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
    dp = P(ctypes.c_double)
    return {
        "cdet_last_error": (ctypes.c_char_p, []),
        "cdet_params_derive": (i, [P(CdetParams)]),
        "cdet_ctx_create": (i, [P(CdetParams), P(vp)]),
        "cdet_ctx_destroy": (i, [vp]),
        "cdet_ctx_stream": (vp, [vp]),
        "cdet_stream_create": (i, [vp, P(vp)]),
        "cdet_stream_destroy": (i, [vp, vp]),
        "cdet_tables_upload": (i, [vp, i, dp, dp, dp, dp]),
        "cdet_tables_alias": (i, [vp, i, i]),
        "cdet_tables_build": (i, [vp, P(CdetLattice)]),
        "cdet_tables_download": (i, [vp, i, dp, dp, dp, dp]),
        "cdet_tables_download_blocks": (i, [vp, i, P(ctypes.c_int64), i64, dp, dp, dp, dp]),
        "cdet_eval_single": (i, [vp, dp, i, dp]),
        "cdet_eval_all_subsets": (i, [vp, dp, i, dp]),
        "cdet_eval_all_subsets_noalloc": (i, [vp, vp, i, vp, vp, vp]),
        "cdet_batch_plan_create": (i, [vp, P(ctypes.c_int32), i64, i, i, P(vp)]),
        "cdet_batch_plan_destroy": (i, [vp]),
        "cdet_batch_out_size": (i64, [vp]),
        "cdet_batch_vertex_count": (i64, [vp]),
        "cdet_batch_query_count": (i64, [vp]),
        "cdet_batch_contracted": (i, [vp, P(ctypes.c_int64), P(ctypes.c_int64)]),
        "cdet_batch_plan_streams": (i, [vp]),
        "cdet_batch_run": (i, [vp, vp, vp, vp, i, vp]),
        "cdet_batch_plan_reuse_pairs": (i, [vp, i]),
        "cdet_ctx_set_pair_reuse": (i, [vp, i]),
        "cdet_batch_profile": (i, [vp, vp, vp, vp, vp, P(ctypes.c_float)]),
        "cdet_eval_batch": (i, [vp, vp, P(ctypes.c_int32), i64, vp, vp, i, vp]),
        "cdet_green_pair_eval": (i, [vp, vp, vp, i64, vp, i, vp]),
        "cdet_green_pair_profile": (i, [vp, vp, vp, i64, vp, i, vp, P(ctypes.c_float)]),
        "cdet_eval_batch_host": (i, [vp, vp, P(ctypes.c_int32), i64, dp]),
        "cdet_eval_batch_submit": (i, [vp, i, vp, P(ctypes.c_int32), i64]),
        "cdet_eval_batch_wait": (i, [vp, i, dp]),
        "cdet_host_alloc": (vp, [vp, ctypes.c_size_t]),
        "cdet_host_free": (None, [vp, vp]),
        "cdet_ctx_set_batch_mode": (i, [vp, i]),
        "cdet_seed_cond_batch_host": (i, [vp, vp, vp, P(ctypes.c_int32), i64, dp]),
        "cdet_get_counters": (i, [vp, P(CdetCounters)]),
        "cdet_ctx_set_tuning": (i, [vp, P(CdetTuning)]),
        "cdet_ctx_get_tuning": (i, [vp, P(CdetTuning)]),
        "cdet_ctx_set_sum_order": (i, [vp, i]),
        "cdet_reset_counters": (i, [vp]),
        # include/cdet_chunks.h
        "cdet_chunks_write_plan": (i, [ctypes.c_char_p, P(ChunkSpec)]),
        "cdet_chunks_read_spec": (i, [ctypes.c_char_p, P(ChunkSpec)]),
        "cdet_chunks_load_tables": (i, [vp, ctypes.c_char_p, P(CdetLattice)]),
        "cdet_chunks_transform_samples": (i, [vp, ctypes.c_char_p, P(CdetLattice)]),
        "cdet_chunks_write_tables": (i, [vp, ctypes.c_char_p, P(ChunkSpec)]),
        "cdet_chunks_write_samples": (i, [vp, ctypes.c_char_p, P(CdetLattice)]),
    }


def lib():
    """Load libcdet_hip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libcdet_hip.so is not built; run `python cdet-renyi-entanglement-hpc_amd/build.py` "
                "(or __graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _signatures().items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _err():
    return lib().cdet_last_error().decode()


def _check(rc):
    if rc != 0:
        raise RuntimeError(_err())


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_params(N1, N2, cheb, beta, U, mu, alpha_up, alpha_down, branch_epsilon=1e-13, tol_fpm=1e-7,
                fpm_global_shift=-1.0, ztransform_epsilon0=-1.0, ztransform_safety=1e4, derived=None):
    """cdet_params with the derived scalars filled (or taken from `derived`)."""
    p = CdetParams(N1=N1, N2=N2, cheb_nu=cheb, cheb_nv=cheb, beta=beta, U=U, mu=mu,
                   alpha_up=alpha_up, alpha_down=alpha_down, branch_epsilon=branch_epsilon,
                   tol_fpm=tol_fpm, fpm_global_shift=fpm_global_shift,
                   ztransform_epsilon0=ztransform_epsilon0, ztransform_safety=ztransform_safety)
    _check(lib().cdet_params_derive(ctypes.byref(p)))
    if derived is not None:
        p.scale, p.diag_up_shift, p.diag_down_shift = derived
    return p


class Context:
    """One device context: parameters, resident tables, scratch and stream."""

    def __init__(self, params):
        self.params = params
        self._plans = weakref.WeakSet()
        h = ctypes.c_void_p()
        _check(lib().cdet_ctx_create(ctypes.byref(params), ctypes.byref(h)))
        self._h = h

    def close(self):
        """Destroy the context, after closing every plan made on it (a plan
        must not outlive its context)."""
        if getattr(self, "_h", None):
            for plan in list(getattr(self, "_plans", ())):
                plan.close()
            lib().cdet_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter exit the module globals may already be gone
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def stream(self):
        return lib().cdet_ctx_stream(self._h)

    def stream_create(self):
        """A hipStream_t (int) on a hardware queue of its own over every CU
        (include/cdet_hip.h cdet_stream_create): batches run concurrently
        overlap only on such streams.  Destroy with stream_destroy before the
        context closes."""
        s = ctypes.c_void_p()
        _check(lib().cdet_stream_create(self._h, ctypes.byref(s)))
        return int(s.value)

    def stream_destroy(self, stream):
        _check(lib().cdet_stream_destroy(self._h, ctypes.c_void_p(int(stream))))

    # -- tables ------------------------------------------------------------
    def upload_table(self, kind, upper, lower, ucorner, lcorner):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (upper, lower, ucorner, lcorner)]
        _check(lib().cdet_tables_upload(self._h, kind, *[_dptr(a) for a in arrs]))

    def alias_table(self, dst, src):
        _check(lib().cdet_tables_alias(self._h, dst, src))

    def upload_tables(self, tables):
        """tables: four objects with .upper/.lower/.ucorner/.lcorner; identical
        objects are uploaded once and aliased (load_data.h:70-91)."""
        seen = {}
        for kind, t in enumerate(tables):
            if id(t) in seen:
                self.alias_table(kind, seen[id(t)])
            else:
                self.upload_table(kind, t.upper, t.lower, t.ucorner, t.lcorner)
                seen[id(t)] = kind

    def build_tables(self, pbc=True, mask="first_row", mask_param=0, node_chunk=0,
                     ordinary_layout=ORDINARY_RESIDENT):
        lat = lattice(pbc, mask, mask_param, node_chunk, ordinary_layout)
        _check(lib().cdet_tables_build(self._h, ctypes.byref(lat)))

    # -- chunk directories (include/cdet_chunks.h) --------------------------
    def load_chunks(self, directory, expect=None):
        """load_chunks workflow: coeffs/ + corners/ streamed into HBM."""
        _check(lib().cdet_chunks_load_tables(self._h, str(directory).encode(),
                                             ctypes.byref(expect) if expect is not None else None))

    def transform_chunk_samples(self, directory, expect=None):
        """Transform stage on the GPU: samples/ + corners/ into HBM tables."""
        _check(lib().cdet_chunks_transform_samples(self._h, str(directory).encode(),
                                                   ctypes.byref(expect) if expect is not None else None))

    def write_chunks(self, directory, spec):
        """Device tables (chunked ordinary layout) to a chunk directory."""
        _check(lib().cdet_chunks_write_tables(self._h, str(directory).encode(), ctypes.byref(spec)))

    def write_chunk_samples(self, directory, lat):
        """samples/ + corners/ of every node job, by the GPU producer."""
        _check(lib().cdet_chunks_write_samples(self._h, str(directory).encode(), ctypes.byref(lat)))

    def download_table(self, kind):
        p = self.params
        S = p.N1 * p.N2
        SP = 4 * S * S
        nc = SP * p.cheb_nu * p.cheb_nv
        up, lo, uc, lc = np.zeros(nc), np.zeros(nc), np.zeros(SP), np.zeros(SP)
        _check(lib().cdet_tables_download(self._h, kind, _dptr(up), _dptr(lo), _dptr(uc), _dptr(lc)))
        return up, lo, uc, lc

    def download_blocks(self, kind, sitepairs):
        """Coefficient blocks and corners of the given sitepairs of a table
        kind: (upper, lower) as (len, nu, nv) arrays, (ucorner, lcorner)."""
        p = self.params
        sp = np.ascontiguousarray(sitepairs, dtype=np.int64)
        n = len(sp)
        up = np.zeros((n, p.cheb_nu, p.cheb_nv))
        lo = np.zeros((n, p.cheb_nu, p.cheb_nv))
        uc, lc = np.zeros(n), np.zeros(n)
        _check(lib().cdet_tables_download_blocks(self._h, kind, sp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                 n, _dptr(up), _dptr(lo), _dptr(uc), _dptr(lc)))
        return up, lo, uc, lc

    # -- integrand (reference API) ----------------------------------------
    def evaluate_integrand(self, vertices):
        """evaluate_cdet_density_integrand_double (cdet_integrand_evaluator.h:542-560)."""
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 4)
        out = np.zeros(1)
        _check(lib().cdet_eval_single(self._h, _dptr(v), len(v), _dptr(out)))
        return float(out[0])

    def evaluate_all_subset_integrands(self, vertices):
        """evaluate_cdet_density_all_subset_integrands_double (:562-590)."""
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 4)
        out = np.zeros(1 << len(v))
        _check(lib().cdet_eval_all_subsets(self._h, _dptr(v), len(v), _dptr(out)))
        return out

    def counters(self):
        c = CdetCounters()
        _check(lib().cdet_get_counters(self._h, ctypes.byref(c)))
        return {k: getattr(c, k) for k, _ in CdetCounters._fields_}

    def reset_counters(self):
        _check(lib().cdet_reset_counters(self._h))

    # -- launch-shape tuning (include/cdet_hip.h cdet_tuning) ---------------
    def set_tuning(self, tuning):
        """Apply a tuning table (a dict as tune.py writes it, or a
        CdetTuning); plans created afterwards use it."""
        t = tuning if isinstance(tuning, CdetTuning) else CdetTuning.from_dict(tuning)
        _check(lib().cdet_ctx_set_tuning(self._h, ctypes.byref(t)))

    def set_sum_order(self, order):
        """Order of the subset recursion's sums in the non-exact modes:
        "tree" (default) or "reference" (include/cdet_hip.h
        cdet_ctx_set_sum_order); the exact mode always sums in the
        reference's order."""
        _check(lib().cdet_ctx_set_sum_order(self._h, SUM_ORDERS[order]))

    def tuning(self):
        t = CdetTuning()
        _check(lib().cdet_ctx_get_tuning(self._h, ctypes.byref(t)))
        return t.to_dict()

    def plan(self, orders, mode="exact", single=False, kind=None):
        plan = BatchPlan(self, orders, mode, single, kind)
        self._plans.add(plan)
        return plan


class BatchPlan:
    """Order-bucketed batched evaluation over device buffers (torch tensors or
    raw device pointers)."""

    def __init__(self, ctx, orders, mode="exact", single=False, kind=None):
        """kind: 0 all-subset values, 1 single-order values, 2 conditioning
        scales of the all-subset values, 3 first-order error bounds of the
        all-subset values (defaults from `single`)."""
        self.ctx = ctx
        self.orders = np.ascontiguousarray(orders, dtype=np.int32)
        self.mode = MODES[mode] if isinstance(mode, str) else int(mode)
        h = ctypes.c_void_p()
        _check(lib().cdet_batch_plan_create(
            ctx.handle, self.orders.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            len(self.orders), self.mode, (1 if single else 0) if kind is None else kind, ctypes.byref(h)))
        self._h = h
        self.out_size = lib().cdet_batch_out_size(h)
        self.vertex_count = lib().cdet_batch_vertex_count(h)
        self.query_count = lib().cdet_batch_query_count(h)
        self.streams = lib().cdet_batch_plan_streams(h)

    def contracted(self):
        """(lookups, tiles) the last run sent to the table contraction."""
        lk, nt = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().cdet_batch_contracted(self._h, ctypes.byref(lk), ctypes.byref(nt)))
        return lk.value, nt.value

    def close(self):
        if getattr(self, "_h", None):
            lib().cdet_batch_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter exit the module globals may already be gone
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _ptr(x):
        if x is None:
            return None
        if hasattr(x, "data_ptr"):
            return ctypes.c_void_p(x.data_ptr())
        return ctypes.c_void_p(int(x))

    def profile(self, d_vertices, d_out, d_status=None, stream=None):
        """Stage times in ms: (lookup grouping, contraction, minors+recursion)."""
        ms = (ctypes.c_float * 3)()
        _check(lib().cdet_batch_profile(self._h, self._ptr(d_vertices), self._ptr(d_out), self._ptr(d_status),
                                        self._ptr(stream), ms))
        return tuple(ms)

    def reuse_pairs(self, enable=True):
        """Green-pair reuse across this plan's runs (cdet_batch_plan_reuse_pairs)."""
        _check(lib().cdet_batch_plan_reuse_pairs(self._h, 1 if enable else 0))

    def run(self, d_vertices, d_out, d_status=None, use_graph=False, stream=None):
        _check(lib().cdet_batch_run(self._h, self._ptr(d_vertices), self._ptr(d_out), self._ptr(d_status),
                                    1 if use_graph else 0, self._ptr(stream)))


def packed_offsets(orders, single=False):
    """Vertex and output offsets of a packed batch (cdet_batch_plan_create)."""
    orders = np.asarray(orders, dtype=np.int64)
    vb = np.concatenate([[0], np.cumsum(orders)])
    sizes = np.ones_like(orders) if single else (np.int64(1) << orders)
    ob = np.concatenate([[0], np.cumsum(sizes)])
    return vb, ob

"""ctypes bindings of the CPU oracle (test infrastructure only).

The oracle is a CPU restatement of the reference's CDet integrand path and its
FP64 table producer (oracle/cdet_oracle.cpp).  Only tests/, the smoke check and
the CPU-baseline leg of bench.py use it, and only as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class CdetParams(ctypes.Structure):
    """Mirror of cdet_params in include/cdet_hip.h."""

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


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        d, i, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        dp = P(ctypes.c_double)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_params_derive.argtypes = [P(CdetParams)]
        L.orc_ctx_create.restype = vp
        L.orc_ctx_create.argtypes = [P(CdetParams)]
        L.orc_ctx_destroy.argtypes = [vp]
        L.orc_ctx_set_table.argtypes = [vp, i, dp, dp, dp, dp]
        L.orc_ctx_fallbacks.restype = ctypes.c_uint64
        L.orc_ctx_fallbacks.argtypes = [vp]
        L.orc_eval_all_subsets.argtypes = [vp, dp, i, dp]
        L.orc_eval_all_subsets_kind.argtypes = [vp, dp, i, i, dp]
        L.orc_eval_single.argtypes = [vp, dp, i, dp]
        L.orc_vertex_matrices.argtypes = [vp, dp, i, i, dp, dp, dp, dp]
        L.orc_mat2pm.argtypes = [dp, i, d, d, dp]
        L.orc_det_minor.restype = d
        L.orc_det_minor.argtypes = [dp, i, ctypes.c_uint64]
        L.orc_ztransform_all.argtypes = [dp, dp, i, d, d, dp]
        L.orc_ztransform_single.restype = d
        L.orc_ztransform_single.argtypes = [dp, dp, i, d, d]
        L.orc_cheb2d.argtypes = [dp, d, d, i, i, dp]
        L.orc_descriptor.argtypes = [i, i, d, d, i, i, i, i, d, d, i, i, i, i, i,
                                     P(ctypes.c_int64), dp]
        L.orc_green_pair.argtypes = [vp, i, i, i, i, i, i, i, i, d, d, dp]
        L.orc_calc_create.restype = vp
        L.orc_calc_create.argtypes = [i, i, d, d, i, i, i]
        L.orc_calc_destroy.argtypes = [vp]
        L.orc_calc_eigen.argtypes = [vp, dp, dp]
        L.orc_calc_mixing.argtypes = [vp, i, dp]
        L.orc_sector_matrix.argtypes = [vp, i, d, d, i, dp]
        L.orc_build_table.argtypes = [vp, i, i, i, dp, dp, dp, dp]
        L.orc_build_samples.argtypes = [vp, i, i, dp, dp, dp, dp]
        L.orc_chebyshev_probes.argtypes = [ctypes.c_uint64, i, i, dp]
        L.orc_uniform_stream.argtypes = [ctypes.c_uint64, d, d, ctypes.c_int64, dp]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


MASK_KINDS = {"empty": 0, "checkerboard": 1, "first_row": 2, "first_column": 3,
              "origin_square": 4, "left_columns": 5}


def make_params(N1, N2, nc, beta, U, mu, alpha_up, alpha_down, branch_epsilon=1e-13,
                tol_fpm=1e-7, fpm_global_shift=-1.0, eps0=-1.0, safety=1e4):
    p = CdetParams(N1=N1, N2=N2, cheb_nu=nc, cheb_nv=nc, beta=beta, U=U, mu=mu,
                   alpha_up=alpha_up, alpha_down=alpha_down, branch_epsilon=branch_epsilon,
                   tol_fpm=tol_fpm, fpm_global_shift=fpm_global_shift,
                   ztransform_epsilon0=eps0, ztransform_safety=safety)
    lib().orc_params_derive(ctypes.byref(p))
    return p


class Table:
    """One resident table kind: upper/lower coefficients and corners."""

    def __init__(self, upper, lower, ucorner, lcorner):
        self.upper = np.ascontiguousarray(upper, dtype=np.float64)
        self.lower = np.ascontiguousarray(lower, dtype=np.float64)
        self.ucorner = np.ascontiguousarray(ucorner, dtype=np.float64)
        self.lcorner = np.ascontiguousarray(lcorner, dtype=np.float64)


class Calc:
    """Free graded two-replica Green calculator (replica_free_greens.h:171-701)."""

    def __init__(self, N1, N2, beta, mu, pbc=True, mask="first_row", mask_param=0):
        self.N1, self.N2, self.S = N1, N2, N1 * N2
        kind = MASK_KINDS[mask] if isinstance(mask, str) else int(mask)
        self._h = lib().orc_calc_create(N1, N2, beta, mu, int(pbc), kind, mask_param)
        if not self._h:
            raise RuntimeError(lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_calc_destroy(self._h)
            self._h = None

    def eigen(self):
        ev = np.zeros(self.S)
        V = np.zeros((self.S, self.S))
        lib().orc_calc_eigen(self._h, _ptr(ev), _ptr(V))
        return ev, V

    def mixing(self, m):
        out = np.zeros((self.S, self.S, 2))
        lib().orc_calc_mixing(self._h, m, _ptr(out))
        return out[..., 0] + 1j * out[..., 1]

    def sector_matrix(self, m, tau, taup, upper):
        out = np.zeros((self.S, self.S, 2))
        _check(lib().orc_sector_matrix(self._h, m, tau, taup, int(upper), _ptr(out)))
        return out[..., 0] + 1j * out[..., 1]

    def build_table(self, nc, ordinary=False, nv=None):
        nv = nc if nv is None else nv
        SP = 4 * self.S * self.S
        up = np.zeros(SP * nc * nv)
        lo = np.zeros(SP * nc * nv)
        uc = np.zeros(SP)
        lc = np.zeros(SP)
        _check(lib().orc_build_table(self._h, nc, nv, int(ordinary), _ptr(up), _ptr(lo), _ptr(uc), _ptr(lc)))
        return Table(up, lo, uc, lc)

    def build_samples(self, nu, nv):
        """Untransformed node samples [sitepair][node] of all replica blocks
        (the chunked workflow's sample stage) and the corner values."""
        SP = 4 * self.S * self.S
        up = np.zeros(SP * nu * nv)
        lo = np.zeros(SP * nu * nv)
        uc = np.zeros(SP)
        lc = np.zeros(SP)
        _check(lib().orc_build_samples(self._h, nu, nv, _ptr(up), _ptr(lo), _ptr(uc), _ptr(lc)))
        return Table(up, lo, uc, lc)


def build_tables(N1, N2, beta, mu, alpha_up, alpha_down, nc, pbc=True, mask="first_row", mask_param=0):
    """Four resident table kinds, as load_data_into_resident_tables (load_data.h:49-99)."""
    tw_up = Calc(N1, N2, beta, mu - alpha_up, pbc, mask, mask_param).build_table(nc)
    od_up = Calc(N1, N2, beta, mu - alpha_up, pbc, "empty", 0).build_table(nc, ordinary=True)
    if alpha_up == alpha_down:
        return [tw_up, tw_up, od_up, od_up]
    tw_dn = Calc(N1, N2, beta, mu - alpha_down, pbc, mask, mask_param).build_table(nc)
    od_dn = Calc(N1, N2, beta, mu - alpha_down, pbc, "empty", 0).build_table(nc, ordinary=True)
    return [tw_up, tw_dn, od_up, od_dn]


class Oracle:
    """Integrand evaluator over host tables (cdet_integrand_evaluator.h)."""

    def __init__(self, params, tables):
        self.params = params
        self.tables = tables
        self._h = lib().orc_ctx_create(ctypes.byref(params))
        for k, t in enumerate(tables):
            lib().orc_ctx_set_table(self._h, k, _ptr(t.upper), _ptr(t.lower), _ptr(t.ucorner), _ptr(t.lcorner))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_ctx_destroy(self._h)
            self._h = None

    @property
    def fpm_fallbacks(self):
        return lib().orc_ctx_fallbacks(self._h)

    def all_subsets(self, verts):
        v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 4)
        out = np.zeros(1 << len(v))
        _check(lib().orc_eval_all_subsets(self._h, _ptr(v), len(v), _ptr(out)))
        return out

    def all_subsets_kind(self, verts, which):
        v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 4)
        out = np.zeros(1 << len(v))
        _check(lib().orc_eval_all_subsets_kind(self._h, _ptr(v), len(v), which, _ptr(out)))
        return out

    def single(self, verts):
        v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 4)
        out = np.zeros(1)
        _check(lib().orc_eval_single(self._h, _ptr(v), len(v), _ptr(out)))
        return float(out[0])

    def vertex_matrices(self, verts, which):
        v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 4)
        n = len(v)
        mu, md, du, dd = np.zeros((n, n)), np.zeros((n, n)), np.zeros(n), np.zeros(n)
        lib().orc_vertex_matrices(self._h, _ptr(v), n, which, _ptr(mu), _ptr(md), _ptr(du), _ptr(dd))
        return mu, md, du, dd

    def green_pair(self, kind_up, kind_dn, ra, rb, x1, y1, x2, y2, tau1, tau2):
        out = np.zeros(2)
        lib().orc_green_pair(self._h, kind_up, kind_dn, ra, rb, x1, y1, x2, y2, tau1, tau2, _ptr(out))
        return out


def mat2pm(A, tol=0.0, shift=-1.0):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    pm = np.zeros((1 << n) - 1)
    fb = lib().orc_mat2pm(_ptr(A), n, tol, shift, _ptr(pm))
    return pm, fb


def det_minors(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    return np.array([lib().orc_det_minor(_ptr(A), n, m) for m in range(1, 1 << n)])


def ztransform_all(up, dn, n, eps0=0.0, safety=0.0):
    up = np.ascontiguousarray(up, dtype=np.float64)
    dn = np.ascontiguousarray(dn, dtype=np.float64)
    dp = np.zeros(1 << n)
    lib().orc_ztransform_all(_ptr(up), _ptr(dn), n, eps0, safety, _ptr(dp))
    return dp


def ztransform_single(up, dn, n, eps0=0.0, safety=0.0):
    up = np.ascontiguousarray(up, dtype=np.float64)
    dn = np.ascontiguousarray(dn, dtype=np.float64)
    return lib().orc_ztransform_single(_ptr(up), _ptr(dn), n, eps0, safety)


def cheb2d(coeffs, u, v):
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    out = np.zeros(1)
    lib().orc_cheb2d(_ptr(c), u, v, c.shape[0], c.shape[1], _ptr(out))
    return float(out[0])


def descriptor(ra, rb, tau1, tau2, x1, y1, x2, y2, beta, eps, nu, nv, N1, N2, upper):
    oi = np.zeros(4, dtype=np.int64)
    od = np.zeros(2)
    lib().orc_descriptor(ra, rb, tau1, tau2, x1, y1, x2, y2, beta, eps, nu, nv, N1, N2, int(upper),
                         oi.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _ptr(od))
    return dict(upper=bool(oi[0]), corner=bool(oi[1]), corner_index=int(oi[2]), base=int(oi[3]),
                u=float(od[0]), v=float(od[1]))


def uniform_stream(seed, lo, hi, count):
    """mt19937_64 + std::uniform_real_distribution(lo, hi), libstdc++ semantics."""
    out = np.zeros(count)
    lib().orc_uniform_stream(seed, lo, hi, count, _ptr(out))
    return out


def chebyshev_probes(seed, count, coord_max):
    out = np.zeros((count, 8))
    lib().orc_chebyshev_probes(seed, count, coord_max, _ptr(out))
    return out

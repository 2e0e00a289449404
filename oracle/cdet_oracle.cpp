// cdet_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement, in plain IEEE double, of the reference's CDet integrand
// path and of its FP64 table producer.  It is the checker of the HIP library:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// it, never the product path.  Every function cites the reference file:line
// whose arithmetic it follows; expressions keep the reference's evaluation
// order (no FMA contraction: built with -ffp-contract=off; explicit fma only
// where the reference calls std::fma).
//
// Pinning (see tests/test_oracle_*.py):
//   clenshaw / 2D table eval .... direct sum of c_ab T_a(u) T_b(v) in long
//                                 double; unit-coefficient known answers
//   mat2pm ...................... brute-force pivoted determinants of every
//                                 principal submatrix, the reference's seeded
//                                 matrices and its pivot-pathology fallback case
//   ztransform (both) ........... independent set-partition cumulant; the
//                                 constant-moment cancellation; single-anchor
//                                 == all-subset[full]
//   all-subset vs single-order .. the reference's 2x2 integrand test
//   table producer .............. direct (unscaled) kernel formula, equal-time
//                                 jump, sector conjugacy, graded boundary
//                                 condition, beta=200 finiteness, Chebyshev
//                                 accuracy and corner tables
//   prefactors / DD combine ..... exact rationals
// The absolute integrand values of the reference binary are not available
// (the reference needs Boost.Multiprecision, absent here): that level is
// parity unpinned and stated so in the design notes.

#include "../include/cdet_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include <quadmath.h>

namespace orc {

// ---------------------------------------------------------------------------
// Double-double arithmetic (include/cdet/utility/types.h:30-150).
// Standard Dekker/Knuth error-free transforms; same operation sequence.
// ---------------------------------------------------------------------------
struct DD {
    double hi = 0.0;
    double lo = 0.0;
};

static inline double quick_two_sum(double a, double b, double& err) {
    const double s = a + b;
    err = b - (s - a);
    return s;
}
static inline double two_sum(double a, double b, double& err) {
    const double s = a + b;
    const double bb = s - a;
    err = (a - (s - bb)) + (b - bb);
    return s;
}
static inline double two_prod(double a, double b, double& err) {
    const double p = a * b;
    err = std::fma(a, b, -p);
    return p;
}
static inline DD dd_add(const DD& x, const DD& y) {  // types.h:72-85
    double s2, t2;
    double s1 = two_sum(x.hi, y.hi, s2);
    const double t1 = two_sum(x.lo, y.lo, t2);
    s2 += t1;
    s1 = quick_two_sum(s1, s2, s2);
    s2 += t2;
    DD r;
    r.hi = quick_two_sum(s1, s2, r.lo);
    return r;
}
static inline DD dd_sub(const DD& x, const DD& y) {  // types.h:87-100
    double s2, t2;
    double s1 = two_sum(x.hi, -y.hi, s2);
    const double t1 = two_sum(x.lo, -y.lo, t2);
    s2 += t1;
    s1 = quick_two_sum(s1, s2, s2);
    s2 += t2;
    DD r;
    r.hi = quick_two_sum(s1, s2, r.lo);
    return r;
}
static inline DD dd_mul(const DD& x, const DD& y) {  // types.h:102-110
    double e;
    const double p = two_prod(x.hi, y.hi, e);
    e += x.hi * y.lo + x.lo * y.hi + x.lo * y.lo;
    DD r;
    r.hi = quick_two_sum(p, e, r.lo);
    return r;
}
static inline DD dd_div(const DD& x, const DD& y) {  // types.h:114-121
    const double q1 = x.hi / y.hi;
    const DD r = dd_sub(x, dd_mul(DD{q1, 0.0}, y));
    const double q2 = (r.hi + r.lo) / y.hi;
    DD q = dd_add(DD{q1, 0.0}, DD{q2, 0.0});
    const double s = q.hi + q.lo;  // renormalize, types.h:65-69
    q.lo = q.lo - (s - q.hi);
    q.hi = s;
    return q;
}
static inline double dd_to_double(const DD& x) { return x.hi + x.lo; }  // types.h:147-149

// Single-order prefactor (-rc)^k / (k! N1 N2) in double-double
// (include/cdet/integrand_replica_randomly_sampled.h:156-160).  (-rc)^k and
// k! are exact integers for k <= 18, so the Boost pow/factorial of the
// reference produce the same double-double operands.
static DD dd_prefactor(int order, int N1, int N2, int rc) {
    double num = 1.0;
    for (int i = 0; i < order; ++i) num *= -static_cast<double>(rc);
    DD fact{1.0, 0.0};
    for (int i = 1; i <= order; ++i) fact = dd_mul(fact, DD{static_cast<double>(i), 0.0});
    const DD den = dd_mul(dd_mul(fact, DD{static_cast<double>(N1), 0.0}), DD{static_cast<double>(N2), 0.0});
    return dd_div(DD{num, 0.0}, den);
}

// All-subset prefactor in long double, cast to double
// (include/cdet/integrand_replica_randomly_sampled.h:161-169).
static double ld_prefactor(int order, int N1, int N2, int rc) {
    long double numerator = 1.0L;
    for (int i = 0; i < order; ++i) numerator *= -static_cast<long double>(rc);
    long double denominator = static_cast<long double>(N1) * static_cast<long double>(N2);
    for (int i = 2; i <= order; ++i) denominator *= static_cast<long double>(i);
    return static_cast<double>(numerator / denominator);
}

// ---------------------------------------------------------------------------
// Lookup descriptor (include/cdet/g_funcs.h:288-342, :609-653;
// include/cdet/replica_chebyshev_tables.h:365-382).
// ---------------------------------------------------------------------------
static inline double clamp_unit(double v) {  // g_funcs.h:300-309
    if (v < -1.0) return -1.0;
    if (v > 1.0) return 1.0;
    return v;
}
static inline int coord_to_index(double coordinate, int extent) {  // g_funcs.h:315-342
    const long double floored = std::floor(static_cast<long double>(coordinate));
    int index = static_cast<int>(floored);
    if (index < 0) index = 0;
    if (index >= extent) index = extent - 1;
    return index;
}
static inline size_t sitepair_index(int ra, int rb, int x1, int y1, int x2, int y2, int N1, int N2) {
    const size_t r = static_cast<size_t>(ra) * 2 + static_cast<size_t>(rb);
    return ((((r * N1 + x1) * N2 + y1) * N1 + x2) * N2) + y2;
}

struct Desc {
    bool upper = false;
    bool corner = false;
    size_t corner_index = 0;
    size_t base = 0;
    double u = 0.0;
    double v = 0.0;
};

static Desc prepare_desc(int ra, int rb, double tau1, double tau2, int x1, int y1, int x2, int y2, double beta,
                         double branch_eps, int nu, int nv, int N1, int N2, bool upper) {
    Desc d;
    d.upper = upper;
    if (upper) {
        if (tau1 < branch_eps) {
            d.corner = true;
            d.corner_index = sitepair_index(ra, rb, x1, y1, x2, y2, N1, N2);
            return d;
        }
        d.u = clamp_unit(2.0 * tau1 / beta - 1.0);
        d.v = clamp_unit(2.0 * tau2 / tau1 - 1.0);
    } else {
        if (tau2 < branch_eps) {
            d.corner = true;
            d.corner_index = sitepair_index(ra, rb, x1, y1, x2, y2, N1, N2);
            return d;
        }
        d.u = clamp_unit(2.0 * tau2 / beta - 1.0);
        d.v = clamp_unit(2.0 * tau1 / tau2 - 1.0);
    }
    d.base = sitepair_index(ra, rb, x1, y1, x2, y2, N1, N2) * static_cast<size_t>(nu) * static_cast<size_t>(nv);
    return d;
}

// Clenshaw pair over contiguous coefficients (include/cdet/g_funcs.h:426-463).
static inline void clenshaw_pair(const double* cu, const double* cd, int degree, double x, double& ou, double& od) {
    if (degree < 0) {
        ou = 0.0;
        od = 0.0;
        return;
    }
    if (degree == 0) {
        ou = cu[0];
        od = cd[0];
        return;
    }
    if (degree == 1) {
        ou = cu[0] + cu[1] * x;
        od = cd[0] + cd[1] * x;
        return;
    }
    const double two_x = 2.0 * x;
    double u1 = 0.0, u2 = 0.0, d1 = 0.0, d2 = 0.0;
    for (int k = degree; k >= 1; --k) {
        const double ub = cu[k] + two_x * u1 - u2;
        const double db = cd[k] + two_x * d1 - d2;
        u2 = u1;
        u1 = ub;
        d2 = d1;
        d1 = db;
    }
    ou = cu[0] + u1 * x - u2;
    od = cd[0] + d1 * x - d2;
}

// 2D sum: rows over v then one Clenshaw over u (include/cdet/g_funcs.h:687-763).
// The fixed-size 25/30/32 paths (:561-578, :502-536) run the same recurrence.
static void cheb2d_pair(const double* cu, const double* cd, double u, double v, int nu, int nv, double& ou,
                        double& od) {
    std::vector<double> ru(static_cast<size_t>(nu)), rd(static_cast<size_t>(nu));
    for (int a = 0; a < nu; ++a) {
        const size_t off = static_cast<size_t>(a) * static_cast<size_t>(nv);
        clenshaw_pair(cu + off, cd + off, nv - 1, v, ru[static_cast<size_t>(a)], rd[static_cast<size_t>(a)]);
    }
    clenshaw_pair(ru.data(), rd.data(), nu - 1, u, ou, od);
}

struct TableView {  // PiecewiseChebyshevTable (replica_chebyshev_tables.h:25-33)
    const double* upper_coeffs = nullptr;
    const double* lower_coeffs = nullptr;
    const double* upper_corners = nullptr;
    const double* lower_corners = nullptr;
};

// include/cdet/g_funcs.h:793-818
static void eval_pair(const Desc& d, const TableView& tu, const TableView& td, int nu, int nv, double& gu,
                      double& gd) {
    if (d.corner) {
        if (d.upper) {
            gu = tu.upper_corners[d.corner_index];
            gd = td.upper_corners[d.corner_index];
        } else {
            gu = tu.lower_corners[d.corner_index];
            gd = td.lower_corners[d.corner_index];
        }
        return;
    }
    if (d.upper) {
        cheb2d_pair(tu.upper_coeffs + d.base, td.upper_coeffs + d.base, d.u, d.v, nu, nv, gu, gd);
    } else {
        cheb2d_pair(tu.lower_coeffs + d.base, td.lower_coeffs + d.base, d.u, d.v, nu, nv, gu, gd);
    }
}

// ---------------------------------------------------------------------------
// Fast principal minors (include/cdet/fast_principle_minors.h).
// ---------------------------------------------------------------------------
static double det_minor_brute(const double* A, int n, size_t mask) {  // :39-105
    std::vector<int> idx;
    for (int i = 0; i < n; ++i)
        if (mask & (size_t(1) << i)) idx.push_back(i);
    const int k = static_cast<int>(idx.size());
    if (k == 0) return 1.0;
    std::vector<double> m(static_cast<size_t>(k * k));
    for (int r = 0; r < k; ++r)
        for (int c = 0; c < k; ++c) m[static_cast<size_t>(r * k + c)] = A[static_cast<size_t>(idx[r] * n + idx[c])];
    double det = 1.0;
    int sign = 1;
    for (int col = 0; col < k; ++col) {
        int prow = col;
        double pabs = std::abs(m[static_cast<size_t>(col * k + col)]);
        for (int r = col + 1; r < k; ++r) {
            const double cand = std::abs(m[static_cast<size_t>(r * k + col)]);
            if (cand > pabs) {
                pabs = cand;
                prow = r;
            }
        }
        if (pabs == 0.0) return 0.0;
        if (prow != col) {
            for (int c = col; c < k; ++c) std::swap(m[static_cast<size_t>(col * k + c)], m[static_cast<size_t>(prow * k + c)]);
            sign = -sign;
        }
        const double piv = m[static_cast<size_t>(col * k + col)];
        det *= piv;
        for (int r = col + 1; r < k; ++r) {
            const double f = m[static_cast<size_t>(r * k + col)] / piv;
            if (f == 0.0) continue;
            m[static_cast<size_t>(r * k + col)] = 0.0;
            for (int c = col + 1; c < k; ++c)
                m[static_cast<size_t>(r * k + c)] = std::fma(-f, m[static_cast<size_t>(col * k + c)], m[static_cast<size_t>(r * k + c)]);
        }
    }
    return sign < 0 ? -det : det;
}

// All principal minors; pm has 2^n - 1 entries, pm[mask-1] = det(A[mask,mask]).
// Returns 1 when the brute-force fallback was taken (:180-186).
static int mat2pm(const double* A, int n, double tol, double thresh, double* pm) {  // :120-246
    const size_t subset_count = size_t(1) << n;
    for (size_t i = 0; i + 1 < subset_count; ++i) pm[i] = 0.0;
    double max_abs = 0.0;
    const int msize = n * n;
    for (int i = 0; i < msize; ++i) {
        const double e = std::abs(A[i]);
        if (e > max_abs) max_abs = e;
    }
    const bool explicit_shift = !(thresh < 0.0);
    const double p = explicit_shift ? thresh : max_abs / static_cast<double>(msize);
    if (max_abs == 0.0 && p == 0.0) return 0;
    const double fb_eps = (tol > 0.0 && max_abs > 0.0) ? tol * max_abs : 0.0;

    size_t max_list = static_cast<size_t>(msize);
    for (int l = 0; l < n - 1; ++l) {
        const size_t ns = static_cast<size_t>(n - l - 1);
        max_list = std::max(max_list, (size_t(1) << (l + 1)) * ns * ns);
    }
    std::vector<double> oldl(max_list), newl(max_list);
    std::copy_n(A, msize, oldl.data());
    std::vector<double> sm(subset_count, p);
    sm[0] = 1.0;
    sm[1] += oldl[0];
    size_t minor_index = 2;
    for (int l = 0; l < n - 1; ++l) {
        const size_t os = static_cast<size_t>(n - l);
        const size_t oms = os * os;
        const size_t ns = os - 1;
        const size_t nms = ns * ns;
        const size_t sets = size_t(1) << l;
        for (size_t s = 0; s < sets; ++s) {
            size_t dw = s * nms;
            const size_t src = s * oms;
            const size_t first_col = src + os;
            size_t row_start = first_col + 1;
            size_t dwith = dw + sets * nms;
            const double pd = oldl[src] + p;
            const double pda = std::abs(pd);
            if (pda == 0.0 || (fb_eps > 0.0 && pda <= fb_eps)) {
                for (size_t mask = 1; mask < subset_count; ++mask) pm[mask - 1] = det_minor_brute(A, n, mask);
                return 1;
            }
            const double inv = 1.0 / pd;
            size_t col_start = first_col;
            for (size_t r = 0; r < ns; ++r) {
                size_t top = src + 1;
                const double f = oldl[col_start] * inv;
                for (size_t c = 0; c < ns; ++c) {
                    const double base = oldl[row_start];
                    newl[dw] = base;
                    newl[dwith] = std::fma(-oldl[top], f, base);
                    ++top;
                    ++row_start;
                    ++dw;
                    ++dwith;
                }
                ++row_start;
                col_start += os;
            }
        }
        for (size_t m = 0; m < 2 * sets; ++m) sm[minor_index++] += newl[m * nms];
        oldl.swap(newl);
    }
    size_t step = 1;
    for (int l = 0; l < n; ++l) {
        for (size_t s = step + 1; s < 2 * step; ++s) sm[s] *= sm[s - step];
        step <<= 1;
    }
    step = subset_count;
    for (int l = n; l > 0; --l) {
        const size_t half = step >> 1;
        for (size_t s = step - 1;; --s) {
            for (size_t t = s; t < subset_count; t += step) sm[t] = std::fma(-p, sm[t - half], sm[t]);
            if (s == half) break;
        }
        step = half;
    }
    for (size_t mask = 1; mask < subset_count; ++mask) pm[mask - 1] = sm[mask];
    return 0;
}

// ---------------------------------------------------------------------------
// Connected-determinant recursions
// (include/cdet/integrand_replica_randomly_sampled.h:226-347).
// ---------------------------------------------------------------------------
static void ztransform_all(const double* up, const double* dn, int n, double eps0, double safety, double* dp) {
    const size_t sc = size_t(1) << n;
    std::vector<double> a(sc, 0.0);
    for (size_t m = 0; m < sc; ++m) dp[m] = 0.0;
    const double thr = eps0 * safety;
    const bool filter = thr > 0.0;
    for (size_t m = 1; m < sc; ++m) a[m] = up[m - 1] * dn[m - 1];
    for (int i = 0; i < n; ++i) dp[size_t(1) << i] = a[size_t(1) << i];
    for (int bits = 2; bits <= n; ++bits) {
        for (size_t mask = 0; mask < sc; ++mask) {
            if (__builtin_popcountll(mask) != bits) continue;
            const size_t anchor = mask & (~mask + 1);
            double acc = a[mask];
            double rhs = std::abs(a[mask]);
            for (size_t sub = (mask - 1) & mask; sub; sub = (sub - 1) & mask) {
                if ((sub & anchor) == 0U) continue;
                const double term = dp[sub] * a[mask ^ sub];
                acc = acc - term;
                rhs += std::abs(term);
            }
            if (filter && std::abs(acc) < thr * rhs) acc = 0.0;
            dp[mask] = acc;
        }
    }
}

static double ztransform_single(const double* up, const double* dn, int n, double eps0, double safety) {
    const size_t sc = size_t(1) << n;
    std::vector<double> a(sc, 0.0), dp(sc, 0.0);
    const double thr = eps0 * safety;
    const bool filter = thr > 0.0;
    for (size_t m = 1; m < sc; ++m) a[m] = up[m - 1] * dn[m - 1];
    for (int i = 0; i < n; ++i) dp[size_t(1) << i] = a[size_t(1) << i];
    for (int bits = 2; bits <= n; ++bits) {
        for (size_t mask = 0; mask < sc; ++mask) {
            if (__builtin_popcountll(mask) != bits) continue;
            double acc = a[mask];
            double rhs = std::abs(a[mask]);
            if (mask & 1U) {
                for (size_t sub = (mask - 1) & mask; sub; sub = (sub - 1) & mask) {
                    if ((sub & 1U) == 0U) continue;
                    const double term = dp[sub] * a[mask ^ sub];
                    acc = acc - term;
                    rhs += std::abs(term);
                }
                if (filter && std::abs(acc) < thr * rhs) acc = 0.0;
            }
            dp[mask] = acc;
        }
    }
    return dp[sc - 1];
}

// ---------------------------------------------------------------------------
// Integrand (include/cdet/integrand_replica_randomly_sampled.h:349-878,
// include/cdet/cdet_integrand_evaluator.h:144-351).
// ---------------------------------------------------------------------------
struct Ctx {
    cdet_params p{};
    TableView tab[4];
    uint64_t fpm_fallbacks = 0;
};

struct Sampled {
    double tau = 0.0;
    double tau_unit = 0.0;
    int x = 0;
    int y = 0;
    int replica = 0;
};

static Sampled decode_vertex(const cdet_vertex& v, const cdet_params& p, int rc) {  // :652-662
    Sampled s;
    s.tau_unit = v.tau;
    s.tau = p.beta * s.tau_unit;
    s.x = coord_to_index(static_cast<double>(p.N1) * v.x, p.N1);
    s.y = coord_to_index(static_cast<double>(p.N2) * v.y, p.N2);
    if (rc <= 1) {
        s.replica = 0;
    } else {
        const int r = coord_to_index(static_cast<double>(rc) * v.replica, rc);  // :142-153
        s.replica = std::clamp(r, 0, rc - 1);
    }
    return s;
}

static double eps0_of(const cdet_params& p) { return p.ztransform_epsilon0 > 0.0 ? p.ztransform_epsilon0 : 2.0e-16; }

// Fill the scaled vertex matrices and unscaled shifted diagonals (:704-776).
static void fill_pairs(Ctx& c, const Sampled* vs, int n, const TableView& tu, const TableView& td, double* m_up,
                       double* m_dn, double* diag_up, double* diag_dn) {
    const cdet_params& p = c.p;
    for (int idx = 0; idx < n * n; ++idx) {
        const int i = idx / n;
        const int j = idx - i * n;
        const bool upper = (i != j) && (vs[i].tau_unit > vs[j].tau_unit);
        const Desc d = prepare_desc(vs[i].replica, vs[j].replica, vs[i].tau, vs[j].tau, vs[i].x, vs[i].y, vs[j].x,
                                    vs[j].y, p.beta, p.branch_epsilon, p.cheb_nu, p.cheb_nv, p.N1, p.N2, upper);
        double gu, gd;
        eval_pair(d, tu, td, p.cheb_nu, p.cheb_nv, gu, gd);
        if (i == j) {
            if (diag_up) diag_up[i] = gu + p.diag_up_shift;
            if (diag_dn) diag_dn[i] = gd + p.diag_down_shift;
            gu += p.diag_up_shift;
            gd += p.diag_down_shift;
        }
        m_up[idx] = gu * p.scale;
        m_dn[idx] = gd * p.scale;
    }
}

// One table class of the all-subset integrand (:549-878); out has 2^n doubles.
static int all_subset_kind(Ctx& c, const cdet_vertex* verts, int n, int kind_up, int kind_dn, int rc, double* out) {
    const cdet_params& p = c.p;
    std::vector<Sampled> vs(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) vs[static_cast<size_t>(i)] = decode_vertex(verts[i], p, rc);
    std::vector<double> mu(static_cast<size_t>(n * n)), md(static_cast<size_t>(n * n)), du(static_cast<size_t>(n)),
        dd(static_cast<size_t>(n));
    fill_pairs(c, vs.data(), n, c.tab[kind_up], c.tab[kind_dn], mu.data(), md.data(), du.data(), dd.data());
    const size_t sc = size_t(1) << n;
    std::vector<double> pu(sc - 1), pd(sc - 1), dp(sc);
    c.fpm_fallbacks += static_cast<uint64_t>(mat2pm(mu.data(), n, p.tol_fpm, p.fpm_global_shift, pu.data()));
    c.fpm_fallbacks += static_cast<uint64_t>(mat2pm(md.data(), n, p.tol_fpm, p.fpm_global_shift, pd.data()));
    ztransform_all(pu.data(), pd.data(), n, eps0_of(p), p.ztransform_safety, dp.data());
    out[0] = 0.0;
    const double counter = p.diag_up_shift * p.diag_down_shift;
    for (int i = 0; i < n; ++i) {
        const double raw = static_cast<double>(-rc) * p.U * p.beta * (du[static_cast<size_t>(i)] * dd[static_cast<size_t>(i)] - counter);
        if (!std::isfinite(raw)) throw std::runtime_error("Error (order-1 singleton integrand): Value is not finite");
        out[size_t(1) << i] = raw;
    }
    double pref[CDET_MAX_ORDER + 2];
    for (int k = 2; k <= n; ++k) pref[k] = ld_prefactor(k, p.N1, p.N2, rc);
    for (size_t mask = 1; mask < sc; ++mask) {
        const int k = __builtin_popcountll(mask);
        if (k <= 1) continue;
        const double v = pref[k] * dp[mask];
        if (!std::isfinite(v)) throw std::runtime_error("Error (all-subset integrand): Value is not finite");
        out[mask] = v;
    }
    return 0;
}

// Single-order integrand of one table class, double-double result (:349-547).
static DD single_kind(Ctx& c, const cdet_vertex* verts, int n, int kind_up, int kind_dn, int rc) {
    const cdet_params& p = c.p;
    std::vector<Sampled> vs(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) vs[static_cast<size_t>(i)] = decode_vertex(verts[i], p, rc);
    std::vector<double> mu(static_cast<size_t>(n * n)), md(static_cast<size_t>(n * n));
    fill_pairs(c, vs.data(), n, c.tab[kind_up], c.tab[kind_dn], mu.data(), md.data(), nullptr, nullptr);
    const size_t sc = size_t(1) << n;
    std::vector<double> pu(sc - 1), pd(sc - 1);
    c.fpm_fallbacks += static_cast<uint64_t>(mat2pm(mu.data(), n, p.tol_fpm, p.fpm_global_shift, pu.data()));
    c.fpm_fallbacks += static_cast<uint64_t>(mat2pm(md.data(), n, p.tol_fpm, p.fpm_global_shift, pd.data()));
    const double cres = ztransform_single(pu.data(), pd.data(), n, eps0_of(p), p.ztransform_safety);
    const DD v = dd_mul(dd_prefactor(n, p.N1, p.N2, rc), DD{cres, 0.0});
    if (!std::isfinite(v.hi + v.lo)) throw std::runtime_error("Error (integrand): Value is not finite");
    return v;
}

// Order-1 value of one table class (:173-213).
static double order1_kind(Ctx& c, const cdet_vertex& vert, int kind_up, int kind_dn, int rc) {
    const cdet_params& p = c.p;
    const Sampled s = decode_vertex(vert, p, rc);
    const Desc d = prepare_desc(s.replica, s.replica, s.tau, s.tau, s.x, s.y, s.x, s.y, p.beta, p.branch_epsilon,
                                p.cheb_nu, p.cheb_nv, p.N1, p.N2, false);
    double gu, gd;
    eval_pair(d, c.tab[kind_up], c.tab[kind_dn], p.cheb_nu, p.cheb_nv, gu, gd);
    const double su = gu + p.diag_up_shift;
    const double sd = gd + p.diag_down_shift;
    const double counter = p.diag_up_shift * p.diag_down_shift;
    return static_cast<double>(-rc) * p.U * p.beta * (su * sd - counter);
}

// ---------------------------------------------------------------------------
// Table producer (include/cdet/replica_free_greens.h, replica_matrix.h,
// fermionic_graded_complex.h, replica_chebyshev_tables.h:432-793).
// ---------------------------------------------------------------------------
struct Cx {
    double re = 0.0;
    double im = 0.0;
};
static inline Cx cadd(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
static inline Cx csub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
static inline Cx cmul(Cx a, Cx b) {  // fermionic_graded_complex.h:37-43
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
static inline Cx cdiv(Cx a, Cx b) {  // :45-56
    const double den = b.re * b.re + b.im * b.im;
    if (den == 0.0) throw std::runtime_error("Division by zero in ComplexNumber");
    return {(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
}
static inline double cabs(Cx a) { return std::sqrt(a.re * a.re + a.im * a.im); }
static inline double real_if_nearly_real(Cx v) {  // :125-133
    const double tol = 1e-12;
    if (std::abs(v.im) > tol * (1.0 + std::abs(v.re)))
        throw std::runtime_error("Complex value expected to be real within tolerance");
    return v.re;
}

static bool site_in_mask(int x, int y, int N1, int N2, int kind, int param) {  // replica_free_greens.h:86-131
    switch (kind) {
        case 0:
            return false;
        case 1:
            return ((x + y) % 2) == 0;
        case 2:
            return y == 0;
        case 3:
            return x == 0;
        case 4: {
            auto in_iv = [&](int c, int extent) {
                if (param <= 0) return false;
                if (2 * param >= extent) return true;
                return c < param || c >= extent - param;
            };
            return in_iv(x, N1) && in_iv(y, N2);
        }
        case 5:
            return x < param;
        default:
            throw std::runtime_error("unknown subregion mask kind");
    }
}

struct Calc {
    int N1 = 0, N2 = 0, S = 0;
    double beta = 0.0, mu = 0.0;
    std::vector<int> mask;
    std::vector<double> evals;
    std::vector<double> V;  // V[site*S + p]
    std::vector<Cx> mix[2];
};

// Jacobi eigensolver (include/cdet/replica_matrix.h:199-300).
static void jacobi(const std::vector<double>& m, int n, std::vector<double>& evals, std::vector<double>& evecs) {
    std::vector<double> a = m, v(static_cast<size_t>(n * n), 0.0);
    for (int i = 0; i < n; ++i) v[static_cast<size_t>(i * n + i)] = 1.0;
    double scale = 0.0;
    for (double x : m) scale = std::max(scale, std::abs(x));
    scale = std::max(scale, 1.0);
    const double tol = 8.0 * std::numeric_limits<double>::epsilon() * scale;
    const int max_sweeps = std::max(20, 20 * n * n);
    auto A = [&](int i, int j) -> double& { return a[static_cast<size_t>(i * n + j)]; };
    auto Vv = [&](int i, int j) -> double& { return v[static_cast<size_t>(i * n + j)]; };
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        double maxo = 0.0;
        int pp = 0, qq = (n > 1) ? 1 : 0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                const double val = std::abs(A(i, j));
                if (val > maxo) {
                    maxo = val;
                    pp = i;
                    qq = j;
                }
            }
        if (n <= 1 || maxo < tol) break;
        const double app = A(pp, pp), aqq = A(qq, qq), apq = A(pp, qq);
        if (apq == 0.0) continue;
        const double tau = (aqq - app) / (2.0 * apq);
        double t;
        if (tau >= 0.0)
            t = 1.0 / (tau + std::sqrt(1.0 + tau * tau));
        else
            t = -1.0 / (std::abs(tau) + std::sqrt(1.0 + tau * tau));
        const double cc = 1.0 / std::sqrt(1.0 + t * t);
        const double s = t * cc;
        for (int k = 0; k < n; ++k) {
            if (k == pp || k == qq) continue;
            const double aik = A(k, pp), akq = A(k, qq);
            A(k, pp) = cc * aik - s * akq;
            A(pp, k) = A(k, pp);
            A(k, qq) = s * aik + cc * akq;
            A(qq, k) = A(k, qq);
        }
        A(pp, pp) = cc * cc * app - 2.0 * s * cc * apq + s * s * aqq;
        A(qq, qq) = s * s * app + 2.0 * s * cc * apq + cc * cc * aqq;
        A(pp, qq) = 0.0;
        A(qq, pp) = 0.0;
        for (int k = 0; k < n; ++k) {
            const double vip = Vv(k, pp), viq = Vv(k, qq);
            Vv(k, pp) = cc * vip - s * viq;
            Vv(k, qq) = s * vip + cc * viq;
        }
        if (sweep + 1 == max_sweeps) throw std::runtime_error("Jacobi diagonalization did not converge");
    }
    std::vector<std::pair<double, int>> ord;
    for (int i = 0; i < n; ++i) ord.emplace_back(A(i, i), i);
    std::sort(ord.begin(), ord.end(), [](const auto& l, const auto& r) { return l.first < r.first; });
    evals.assign(static_cast<size_t>(n), 0.0);
    evecs.assign(static_cast<size_t>(n * n), 0.0);
    for (int nc = 0; nc < n; ++nc) {
        const int oc = ord[static_cast<size_t>(nc)].second;
        evals[static_cast<size_t>(nc)] = ord[static_cast<size_t>(nc)].first;
        for (int r = 0; r < n; ++r) evecs[static_cast<size_t>(r * n + nc)] = Vv(r, oc);
    }
}

static Calc make_calc(int N1, int N2, double beta, double mu, bool pbc, int mask_kind, int mask_param) {
    Calc c;
    c.N1 = N1;
    c.N2 = N2;
    c.S = N1 * N2;
    c.beta = beta;
    c.mu = mu;
    const int S = c.S;
    c.mask.assign(static_cast<size_t>(S), 0);
    for (int y = 0; y < N2; ++y)
        for (int x = 0; x < N1; ++x) c.mask[static_cast<size_t>(y * N1 + x)] = site_in_mask(x, y, N1, N2, mask_kind, mask_param) ? 1 : 0;
    // Square-lattice Hamiltonian (replica_free_greens.h:133-169).
    std::vector<double> h(static_cast<size_t>(S * S), 0.0);
    const double t = 1.0;
    for (int y = 0; y < N2; ++y)
        for (int x = 0; x < N1; ++x) {
            const int i = y * N1 + x;
            h[static_cast<size_t>(i * S + i)] = -mu;
            int nb[2][2];
            int nn = 0;
            if (x + 1 < N1) {
                nb[nn][0] = x + 1;
                nb[nn][1] = y;
                ++nn;
            } else if (pbc) {
                nb[nn][0] = 0;
                nb[nn][1] = y;
                ++nn;
            }
            if (y + 1 < N2) {
                nb[nn][0] = x;
                nb[nn][1] = y + 1;
                ++nn;
            } else if (pbc) {
                nb[nn][0] = x;
                nb[nn][1] = 0;
                ++nn;
            }
            for (int k = 0; k < nn; ++k) {
                const int j = nb[k][1] * N1 + nb[k][0];
                h[static_cast<size_t>(i * S + j)] -= t;
                h[static_cast<size_t>(j * S + i)] -= t;
            }
        }
    jacobi(h, S, c.evals, c.V);
    // Sector mixing S_m = V^T Omega_m^{-1} V (replica_free_greens.h:366-421, :675-684).
    for (int m = 0; m < 2; ++m) {
        c.mix[m].assign(static_cast<size_t>(S * S), Cx{});
        for (int pidx = 0; pidx < S; ++pidx)
            for (int site = 0; site < S; ++site) {
                const double vip = c.V[static_cast<size_t>(site * S + pidx)];
                if (vip == 0.0) continue;
                Cx phase{1.0, 0.0};
                if (c.mask[static_cast<size_t>(site)] == 1) phase = (m == 0) ? Cx{0.0, -1.0} : Cx{0.0, 1.0};
                const Cx wl = cmul(Cx{vip, 0.0}, phase);
                for (int q = 0; q < S; ++q) {
                    Cx& dst = c.mix[m][static_cast<size_t>(pidx * S + q)];
                    dst = cadd(dst, cmul(wl, Cx{c.V[static_cast<size_t>(site * S + q)], 0.0}));
                }
            }
    }
    return c;
}

// Pivoted complex Gauss-Jordan (fermionic_graded_complex.h:304-380).
static void gj_solve(std::vector<Cx>& mat, std::vector<Cx>& rhs, int n) {
    const double tol = 1e-12;
    auto M = [&](int r, int c) -> Cx& { return mat[static_cast<size_t>(r * n + c)]; };
    auto R = [&](int r, int c) -> Cx& { return rhs[static_cast<size_t>(r * n + c)]; };
    for (int col = 0; col < n; ++col) {
        int prow = col;
        double pabs = cabs(M(col, col));
        for (int r = col + 1; r < n; ++r) {
            const double cand = cabs(M(r, col));
            if (cand > pabs) {
                prow = r;
                pabs = cand;
            }
        }
        if (pabs <= tol) throw std::runtime_error("Complex matrix solve is singular within the requested precision");
        if (prow != col) {
            for (int j = 0; j < n; ++j) std::swap(M(col, j), M(prow, j));
            for (int j = 0; j < n; ++j) std::swap(R(col, j), R(prow, j));
        }
        const Cx piv = M(col, col);
        for (int j = col; j < n; ++j) M(col, j) = cdiv(M(col, j), piv);
        for (int j = 0; j < n; ++j) R(col, j) = cdiv(R(col, j), piv);
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const Cx f = M(r, col);
            if (f.re == 0.0 && f.im == 0.0) continue;
            for (int j = col; j < n; ++j) M(r, j) = csub(M(r, j), cmul(f, M(col, j)));
            for (int j = 0; j < n; ++j) R(r, j) = csub(R(r, j), cmul(f, R(col, j)));
        }
    }
}

// Stable scaled sector kernel in the eigenbasis (replica_free_greens.h:423-496).
static std::vector<Cx> sector_kernel(const Calc& c, int m, double tau, double taup, bool upper) {
    const int S = c.S;
    const double tf = tau / c.beta;
    const double tpf = taup / c.beta;
    const double theta = upper ? (tf + tpf) / 2.0 : 1.0 - (tf + tpf) / 2.0;
    const std::vector<Cx>& mix = c.mix[m];
    std::vector<Cx> middle(static_cast<size_t>(S * S)), rhs(static_cast<size_t>(S * S));
    std::vector<double> left(static_cast<size_t>(S)), rs(static_cast<size_t>(S)), cs(static_cast<size_t>(S)),
        dm(static_cast<size_t>(S));
    for (int a = 0; a < S; ++a) {
        const double x = c.beta * c.evals[static_cast<size_t>(a)];
        const double sx = upper ? x : -x;
        const double cc = std::max(0.0, sx);
        double le, re;
        if (upper) {
            le = (1.0 - tf) * x - (1.0 - theta) * cc;
            re = tpf * x - theta * cc;
        } else {
            le = tf * sx - (1.0 - theta) * cc;
            re = (1.0 - tpf) * sx - theta * cc;
        }
        left[static_cast<size_t>(a)] = std::exp(le);
        rhs[static_cast<size_t>(a * S + a)] = Cx{std::exp(re), 0.0};
        rs[static_cast<size_t>(a)] = std::exp(-theta * cc);
        cs[static_cast<size_t>(a)] = std::exp(-(1.0 - theta) * cc);
        dm[static_cast<size_t>(a)] = std::exp(sx - cc);
    }
    for (int a = 0; a < S; ++a) {
        for (int b = 0; b < S; ++b) {
            const Cx mv = upper ? mix[static_cast<size_t>(a * S + b)]
                                : Cx{mix[static_cast<size_t>(b * S + a)].re, -mix[static_cast<size_t>(b * S + a)].im};
            middle[static_cast<size_t>(a * S + b)] =
                cmul(cmul(Cx{rs[static_cast<size_t>(a)], 0.0}, mv), Cx{cs[static_cast<size_t>(b)], 0.0});
        }
        middle[static_cast<size_t>(a * S + a)] = cadd(middle[static_cast<size_t>(a * S + a)], Cx{dm[static_cast<size_t>(a)], 0.0});
    }
    gj_solve(middle, rhs, S);
    for (int a = 0; a < S; ++a) {
        const Cx rsc = upper ? Cx{left[static_cast<size_t>(a)], 0.0} : Cx{-left[static_cast<size_t>(a)], 0.0};
        for (int b = 0; b < S; ++b) rhs[static_cast<size_t>(a * S + b)] = cmul(rsc, rhs[static_cast<size_t>(a * S + b)]);
    }
    return rhs;
}

// G_site = V K V^T (replica_free_greens.h:498-534).
static std::vector<Cx> to_site(const Calc& c, const std::vector<Cx>& K) {
    const int S = c.S;
    std::vector<Cx> tmp(static_cast<size_t>(S * S)), res(static_cast<size_t>(S * S));
    for (int i = 0; i < S; ++i)
        for (int pp = 0; pp < S; ++pp) {
            const double vip = c.V[static_cast<size_t>(i * S + pp)];
            if (vip == 0.0) continue;
            const Cx l{vip, 0.0};
            for (int q = 0; q < S; ++q)
                tmp[static_cast<size_t>(i * S + q)] = cadd(tmp[static_cast<size_t>(i * S + q)], cmul(l, K[static_cast<size_t>(pp * S + q)]));
        }
    for (int i = 0; i < S; ++i)
        for (int q = 0; q < S; ++q) {
            const Cx tiq = tmp[static_cast<size_t>(i * S + q)];
            if (tiq.re == 0.0 && tiq.im == 0.0) continue;
            for (int j = 0; j < S; ++j)
                res[static_cast<size_t>(i * S + j)] = cadd(res[static_cast<size_t>(i * S + j)], cmul(tiq, Cx{c.V[static_cast<size_t>(j * S + q)], 0.0}));
        }
    return res;
}

static std::vector<Cx> sector_matrix(const Calc& c, int m, double tau, double taup, bool upper) {
    return to_site(c, sector_kernel(c, m, tau, taup, upper));
}

// replica_chebyshev_tables.h:489-510
static double replica_entry(const std::vector<Cx>& g0, const std::vector<Cx>& g1, int ra, int rb, int s1, int s2, int S) {
    const Cx a = g0[static_cast<size_t>(s1 * S + s2)];
    const Cx b = g1[static_cast<size_t>(s1 * S + s2)];
    if (ra == rb) return real_if_nearly_real(cdiv(cadd(a, b), Cx{2.0, 0.0}));
    const Cx d = csub(a, b);
    if (ra == 0 && rb == 1) return real_if_nearly_real(cdiv(cmul(Cx{0.0, 1.0}, d), Cx{2.0, 0.0}));
    return real_if_nearly_real(cdiv(cmul(Cx{-0.0, -1.0}, d), Cx{2.0, 0.0}));
}

struct NodeTables {  // replica_chebyshev_tables.h:432-471
    std::vector<double> nu, nv, au, av, cu, cv;
};
static NodeTables node_tables(int nu, int nv) {
    NodeTables t;
    const double pi = M_PI;
    t.nu.resize(static_cast<size_t>(nu));
    t.nv.resize(static_cast<size_t>(nv));
    t.au.resize(static_cast<size_t>(nu));
    t.av.resize(static_cast<size_t>(nv));
    t.cu.resize(static_cast<size_t>(nu * nu));
    t.cv.resize(static_cast<size_t>(nv * nv));
    for (int i = 0; i < nu; ++i) t.nu[static_cast<size_t>(i)] = std::cos((static_cast<double>(2 * i + 1) * pi) / static_cast<double>(2 * nu));
    for (int j = 0; j < nv; ++j) t.nv[static_cast<size_t>(j)] = std::cos((static_cast<double>(2 * j + 1) * pi) / static_cast<double>(2 * nv));
    for (int a = 0; a < nu; ++a) {
        t.au[static_cast<size_t>(a)] = (a == 0) ? 1.0 : 2.0;
        for (int i = 0; i < nu; ++i) {
            const double th = (static_cast<double>(2 * i + 1) * pi) / static_cast<double>(2 * nu);
            t.cu[static_cast<size_t>(a * nu + i)] = std::cos(static_cast<double>(a) * th);
        }
    }
    for (int b = 0; b < nv; ++b) {
        t.av[static_cast<size_t>(b)] = (b == 0) ? 1.0 : 2.0;
        for (int j = 0; j < nv; ++j) {
            const double th = (static_cast<double>(2 * j + 1) * pi) / static_cast<double>(2 * nv);
            t.cv[static_cast<size_t>(b * nv + j)] = std::cos(static_cast<double>(b) * th);
        }
    }
    return t;
}

// replica_chebyshev_tables.h:552-590
static void dct_block(double* block, const NodeTables& t, int nu, int nv) {
    std::vector<double> tmp(static_cast<size_t>(nu * nv), 0.0);
    for (int i = 0; i < nu; ++i) {
        const double* row = block + static_cast<size_t>(i * nv);
        for (int b = 0; b < nv; ++b) {
            const double pv = t.av[static_cast<size_t>(b)] / static_cast<double>(nv);
            const double* cr = t.cv.data() + static_cast<size_t>(b * nv);
            double tot = 0.0;
            for (int j = 0; j < nv; ++j) tot += row[j] * cr[j];
            tmp[static_cast<size_t>(i * nv + b)] = pv * tot;
        }
    }
    for (int a = 0; a < nu; ++a) {
        const double pu = t.au[static_cast<size_t>(a)] / static_cast<double>(nu);
        const double* cr = t.cu.data() + static_cast<size_t>(a * nu);
        for (int b = 0; b < nv; ++b) {
            double tot = 0.0;
            for (int i = 0; i < nu; ++i) tot += tmp[static_cast<size_t>(i * nv + b)] * cr[i];
            block[static_cast<size_t>(a * nv + b)] = pu * tot;
        }
    }
}

// Node times (replica_chebyshev_tables.h:473-487, :529-550).
static void node_time(double cu, double cv, bool upper, double beta, double& tau, double& taup) {
    if (upper) {
        tau = beta * (1.0 + cu) / 2.0;
        taup = tau * (1.0 + cv) / 2.0;
    } else {
        taup = beta * (1.0 + cu) / 2.0;
        tau = taup * (1.0 + cv) / 2.0;
    }
}

// Twisted table (replica_chebyshev_tables.h:592-695), or the ordinary
// replica-(0,0) table in site-major order (:697-793) when ordinary != 0.
// transform = false leaves the node samples untransformed at [sitepair][node]
// (the chunked workflow's sample stage, chunked_load_data.h:947-996).
static void build_table(const Calc& c, int nu, int nv, bool ordinary, double* upper, double* lower, double* ucorner,
                        double* lcorner, bool transform = true) {
    const int S = c.S;
    const size_t SP = static_cast<size_t>(4) * S * S;
    const size_t ncoef = static_cast<size_t>(nu) * nv;
    std::fill(upper, upper + SP * ncoef, 0.0);
    std::fill(lower, lower + SP * ncoef, 0.0);
    std::fill(ucorner, ucorner + SP, 0.0);
    std::fill(lcorner, lcorner + SP, 0.0);
    const NodeTables t = node_tables(nu, nv);
    auto fill_site = [&](const std::vector<Cx>& g0, const std::vector<Cx>& g1, double* dst, size_t stride,
                         size_t offset) {
        if (ordinary) {
            for (int s1 = 0; s1 < S; ++s1)
                for (int s2 = 0; s2 < S; ++s2)
                    dst[static_cast<size_t>(s1 * S + s2) * stride + offset] = real_if_nearly_real(g0[static_cast<size_t>(s1 * S + s2)]);
            return;
        }
        for (size_t flat = 0; flat < SP; ++flat) {
            size_t f = flat;
            const int y2 = static_cast<int>(f % c.N2);
            f /= c.N2;
            const int x2 = static_cast<int>(f % c.N1);
            f /= c.N1;
            const int y1 = static_cast<int>(f % c.N2);
            f /= c.N2;
            const int x1 = static_cast<int>(f % c.N1);
            f /= c.N1;
            const int rb = static_cast<int>(f % 2);
            f /= 2;
            const int ra = static_cast<int>(f % 2);
            dst[flat * stride + offset] = replica_entry(g0, g1, ra, rb, y1 * c.N1 + x1, y2 * c.N1 + x2, S);
        }
    };
    for (int b = 0; b < 2; ++b) {
        const bool up = (b == 0);
        const std::vector<Cx> g0 = sector_matrix(c, 0, 0.0, 0.0, up);
        const std::vector<Cx> g1 = ordinary ? g0 : sector_matrix(c, 1, 0.0, 0.0, up);
        fill_site(g0, g1, up ? ucorner : lcorner, 1, 0);
    }
    for (int b = 0; b < 2; ++b) {
        const bool up = (b == 0);
        double* dst = up ? upper : lower;
        for (int i = 0; i < nu; ++i)
            for (int j = 0; j < nv; ++j) {
                double tau, taup;
                node_time(t.nu[static_cast<size_t>(i)], t.nv[static_cast<size_t>(j)], up, c.beta, tau, taup);
                const std::vector<Cx> g0 = sector_matrix(c, 0, tau, taup, up);
                const std::vector<Cx> g1 = ordinary ? g0 : sector_matrix(c, 1, tau, taup, up);
                fill_site(g0, g1, dst, ncoef, static_cast<size_t>(i * nv + j));
            }
        const size_t active = ordinary ? static_cast<size_t>(S) * S : SP;
        if (transform)
            for (size_t flat = 0; flat < active; ++flat) dct_block(dst + flat * ncoef, t, nu, nv);
    }
}

}  // namespace orc

// ===========================================================================
// extern "C" surface for ctypes (tests / smoke / cpu baseline only).
// ===========================================================================
namespace {
thread_local std::string g_err;
}

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

// Derived scalars, 113-bit arithmetic, rounded once to double.
int orc_params_derive(cdet_params* p) {
    const __float128 prod = static_cast<__float128>(p->U) * static_cast<__float128>(p->N1) *
                            static_cast<__float128>(p->N2) * static_cast<__float128>(p->beta);
    p->scale = static_cast<double>(sqrtq(prod));
    p->diag_up_shift = static_cast<double>(static_cast<__float128>(p->alpha_down) / static_cast<__float128>(p->U));
    p->diag_down_shift = static_cast<double>(static_cast<__float128>(p->alpha_up) / static_cast<__float128>(p->U));
    return 0;
}

void* orc_ctx_create(const cdet_params* p) {
    auto* c = new orc::Ctx();
    c->p = *p;
    return c;
}
void orc_ctx_destroy(void* ctx) { delete static_cast<orc::Ctx*>(ctx); }
int orc_ctx_set_table(void* ctx, int kind, const double* upper, const double* lower, const double* ucorner,
                      const double* lcorner) {
    auto* c = static_cast<orc::Ctx*>(ctx);
    if (kind < 0 || kind > 3) return -1;
    c->tab[kind] = orc::TableView{upper, lower, ucorner, lcorner};
    return 0;
}
uint64_t orc_ctx_fallbacks(void* ctx) { return static_cast<orc::Ctx*>(ctx)->fpm_fallbacks; }

// evaluate_cdet_density_all_subset_integrands_double_into
// (cdet_integrand_evaluator.h:562-583, :304-351).
int orc_eval_all_subsets(void* ctx, const cdet_vertex* v, int E, double* out) {
    try {
        if (E <= 1) throw std::runtime_error("The all-subset replica integrand is defined only for envelope order >= 2");
        auto* c = static_cast<orc::Ctx*>(ctx);
        const size_t sc = size_t(1) << E;
        std::vector<double> ord(sc);
        orc::all_subset_kind(*c, v, E, CDET_TABLE_TWISTED_UP, CDET_TABLE_TWISTED_DOWN, 2, out);
        orc::all_subset_kind(*c, v, E, CDET_TABLE_ORDINARY_UP, CDET_TABLE_ORDINARY_DOWN, 1, ord.data());
        for (size_t m = 0; m < sc; ++m) out[m] -= 2.0 * ord[m];
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// All-subset values of one table class only (kind 0 = twisted rc=2, 1 = ordinary rc=1).
int orc_eval_all_subsets_kind(void* ctx, const cdet_vertex* v, int E, int which, double* out) {
    try {
        auto* c = static_cast<orc::Ctx*>(ctx);
        if (which == 0)
            orc::all_subset_kind(*c, v, E, 0, 1, 2, out);
        else
            orc::all_subset_kind(*c, v, E, 2, 3, 1, out);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// evaluate_cdet_density_integrand_double (cdet_integrand_evaluator.h:542-560).
int orc_eval_single(void* ctx, const cdet_vertex* v, int n, double* out) {
    try {
        if (n <= 0) throw std::runtime_error("The replica integrand is defined only for order >= 1");
        auto* c = static_cast<orc::Ctx*>(ctx);
        if (n == 1) {  // :171-201, combined in DD
            const double tw = orc::order1_kind(*c, v[0], 0, 1, 2);
            const double od = orc::order1_kind(*c, v[0], 2, 3, 1);
            const orc::DD r = orc::dd_sub(orc::DD{tw, 0.0}, orc::dd_mul(orc::DD{2.0, 0.0}, orc::DD{od, 0.0}));
            *out = orc::dd_to_double(r);
            return 0;
        }
        const orc::DD tw = orc::single_kind(*c, v, n, 0, 1, 2);
        const orc::DD od = orc::single_kind(*c, v, n, 2, 3, 1);
        const orc::DD r = orc::dd_sub(tw, orc::dd_mul(orc::DD{2.0, 0.0}, od));
        *out = orc::dd_to_double(r);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Scaled vertex matrices of one table class (for stage-level parity).
int orc_vertex_matrices(void* ctx, const cdet_vertex* v, int n, int which, double* m_up, double* m_dn, double* d_up,
                        double* d_dn) {
    auto* c = static_cast<orc::Ctx*>(ctx);
    const int rc = which == 0 ? 2 : 1;
    std::vector<orc::Sampled> vs(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) vs[static_cast<size_t>(i)] = orc::decode_vertex(v[i], c->p, rc);
    orc::fill_pairs(*c, vs.data(), n, c->tab[which == 0 ? 0 : 2], c->tab[which == 0 ? 1 : 3], m_up, m_dn, d_up, d_dn);
    return 0;
}

int orc_mat2pm(const double* A, int n, double tol, double thresh, double* pm) { return orc::mat2pm(A, n, tol, thresh, pm); }
double orc_det_minor(const double* A, int n, uint64_t mask) { return orc::det_minor_brute(A, n, static_cast<size_t>(mask)); }
void orc_ztransform_all(const double* up, const double* dn, int n, double eps0, double safety, double* dp) {
    orc::ztransform_all(up, dn, n, eps0, safety, dp);
}
double orc_ztransform_single(const double* up, const double* dn, int n, double eps0, double safety) {
    return orc::ztransform_single(up, dn, n, eps0, safety);
}
void orc_cheb2d(const double* coeffs, double u, double v, int nu, int nv, double* out) {
    double a, b;
    orc::cheb2d_pair(coeffs, coeffs, u, v, nu, nv, a, b);
    *out = a;
}
// Descriptor of one lookup: out_i = {upper, corner, corner_index, base}, out_d = {u, v}.
void orc_descriptor(int ra, int rb, double tau1, double tau2, int x1, int y1, int x2, int y2, double beta,
                    double eps, int nu, int nv, int N1, int N2, int upper, int64_t* out_i, double* out_d) {
    const orc::Desc d = orc::prepare_desc(ra, rb, tau1, tau2, x1, y1, x2, y2, beta, eps, nu, nv, N1, N2, upper != 0);
    out_i[0] = d.upper;
    out_i[1] = d.corner;
    out_i[2] = static_cast<int64_t>(d.corner_index);
    out_i[3] = static_cast<int64_t>(d.base);
    out_d[0] = d.u;
    out_d[1] = d.v;
}
// (G_up, G_down) of one lookup through the tables of `kind_up`, `kind_dn`.
void orc_green_pair(void* ctx, int kind_up, int kind_dn, int ra, int rb, int x1, int y1, int x2, int y2, double tau1,
                    double tau2, double* out) {
    auto* c = static_cast<orc::Ctx*>(ctx);
    const cdet_params& p = c->p;
    const orc::Desc d = orc::prepare_desc(ra, rb, tau1, tau2, x1, y1, x2, y2, p.beta, p.branch_epsilon, p.cheb_nu,
                                          p.cheb_nv, p.N1, p.N2, tau1 > tau2);
    orc::eval_pair(d, c->tab[kind_up], c->tab[kind_dn], p.cheb_nu, p.cheb_nv, out[0], out[1]);
}

// ---- producer ----
void* orc_calc_create(int N1, int N2, double beta, double mu, int pbc, int mask_kind, int mask_param) {
    try {
        return new orc::Calc(orc::make_calc(N1, N2, beta, mu, pbc != 0, mask_kind, mask_param));
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void orc_calc_destroy(void* c) { delete static_cast<orc::Calc*>(c); }
void orc_calc_eigen(void* c, double* evals, double* evecs) {
    auto* k = static_cast<orc::Calc*>(c);
    std::copy(k->evals.begin(), k->evals.end(), evals);
    std::copy(k->V.begin(), k->V.end(), evecs);
}
void orc_calc_mixing(void* c, int m, double* out /* S*S*2 */) {
    auto* k = static_cast<orc::Calc*>(c);
    for (size_t i = 0; i < k->mix[m].size(); ++i) {
        out[2 * i] = k->mix[m][i].re;
        out[2 * i + 1] = k->mix[m][i].im;
    }
}
// Sector matrix G_m(tau, tau') in site basis, complex interleaved S*S*2.
int orc_sector_matrix(void* c, int m, double tau, double taup, int upper, double* out) {
    try {
        auto* k = static_cast<orc::Calc*>(c);
        const auto g = orc::sector_matrix(*k, m, tau, taup, upper != 0);
        for (size_t i = 0; i < g.size(); ++i) {
            out[2 * i] = g[i].re;
            out[2 * i + 1] = g[i].im;
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}
int orc_build_table(void* c, int nu, int nv, int ordinary, double* upper, double* lower, double* ucorner,
                    double* lcorner) {
    try {
        orc::build_table(*static_cast<orc::Calc*>(c), nu, nv, ordinary != 0, upper, lower, ucorner, lcorner);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}
// Node samples [sitepair][node] of all replica blocks, before the transform.
int orc_build_samples(void* c, int nu, int nv, double* upper, double* lower, double* ucorner, double* lcorner) {
    try {
        orc::build_table(*static_cast<orc::Calc*>(c), nu, nv, false, upper, lower, ucorner, lcorner, false);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// mt19937_64 + uniform_real_distribution(lo, hi) stream, libstdc++ semantics
// (the reference's tests and sampler draw their inputs this way).
void orc_uniform_stream(uint64_t seed, double lo, double hi, int64_t count, double* out) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> dist(lo, hi);
    for (int64_t i = 0; i < count; ++i) out[i] = dist(rng);
}

// Batch callback of the MCMCMC driver over the oracle (packed envelopes).
int orc_eval_batch_host(void* ctx, const cdet_vertex* v, const int32_t* orders, int64_t count, double* out) {
    int64_t vo = 0, oo = 0;
    for (int64_t c = 0; c < count; ++c) {
        const int rc = orc_eval_all_subsets(ctx, v + vo, orders[c], out + oo);
        if (rc) return rc;
        vo += orders[c];
        oo += int64_t(1) << orders[c];
    }
    return 0;
}

// Synthetic product integrand of the reference's driver test
// (tests/test_mcmcmc_synthetic.cpp:18-28): values[mask] = prod (0.5 + tau_i),
// whose integral is 1 at every order.
int orc_synthetic_batch_eval(void*, const cdet_vertex* v, const int32_t* orders, int64_t count, double* out) {
    int64_t vo = 0, oo = 0;
    for (int64_t c = 0; c < count; ++c) {
        const int n = orders[c];
        const uint64_t sc = uint64_t(1) << n;
        double* vals = out + oo;
        vals[0] = 1.0;
        for (uint64_t m = 1; m < sc; ++m) {
            const uint64_t bit = m & (~m + 1U);
            const int k = __builtin_ctzll(bit);
            vals[m] = vals[m ^ bit] * (0.5 + v[vo + k].tau);
        }
        vo += n;
        oo += static_cast<int64_t>(sc);
    }
    return 0;
}

// Probe set of the reference's Chebyshev accuracy test
// (tests/test_chebyshev_accuracy.cpp:132-152): out[8*k..] = tau1/beta,
// tau2/beta, replica1, replica2, x1, y1, x2, y2.
void orc_chebyshev_probes(uint64_t seed, int count, int coord_max, double* out) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> unit(0.02, 0.98);
    std::uniform_int_distribution<int> replica(0, 1);
    std::uniform_int_distribution<int> coordinate(0, coord_max);
    for (int s = 0; s < count; ++s) {
        double tau1 = unit(rng);
        double tau2 = unit(rng);
        if (std::abs(tau1 - tau2) < 0.03) tau2 = std::fmod(tau2 + 0.17, 0.96) + 0.02;
        double* o = out + 8 * s;
        o[0] = tau1;
        o[1] = tau2;
        o[2] = replica(rng);
        o[3] = replica(rng);
        o[4] = coordinate(rng);
        o[5] = coordinate(rng);
        o[6] = coordinate(rng);
        o[7] = coordinate(rng);
    }
}

}  // extern "C"

// cdet_bound.hpp -- first-order error bound of every all-subset integrand
// value (evaluation kind 3 of a batch plan).
//
// The values cancel heavily: a principal minor is recovered from shifted
// minors by inclusion-exclusion, and a connected value from products of
// minors by the subset recursion.  A blanket relative error bound therefore
// cannot hold for every value, whatever the arithmetic.  This kind computes,
// for each value, a bound on its change under
//   (1) an absolute error of each table value G_ij of at most
//       err_scale * sum_ab |c_ab| over the pair's coefficient block (the
//       summation error of any evaluation order of the Chebyshev sum:
//       |T_k| <= 1 on [-1, 1]; err_scale = (nu + nv) u), and
//   (2) one rounding error u |result| per floating-point operation,
// by running error analysis (Wilkinson): every intermediate x of the
// reference's algorithm carries E(x) >= its first-order perturbation,
// propagated as E(a b) = |a| E(b) + |b| E(a) + u |a b|, E(a +- b) = E(a) +
// E(b) + u |a +- b|, E(1/a) = E(a) / a^2 + u |1/a|.  The algorithm is the
// reference's: the shifted Schur-complement tree of mat2pm with prefix
// products and shift removal (include/cdet/fast_principle_minors.h:120-246),
// its pivoted-determinant fallback (:39-115), the all-subset recursion with
// the round-off filter (include/cdet/integrand_replica_randomly_sampled.h
// :292-347) and the assembly (:833-877).  Two evaluations of the same
// configuration whose table values are both within (1) differ by at most the
// sum of their bounds, about twice the bound either computes.
//
// A filtered value (|acc| < thr * sum|terms|) that another evaluation might
// not filter, or an unfiltered one it might, is bounded by the filter window
// thr * (sum|terms| + E) + E, the most either value can be.
//
// One workgroup per (configuration, class); the work lives in global scratch
// (bound_doubles per workgroup) and the arithmetic order is free, since only
// magnitudes enter the bound.  This kind is for error analysis, not speed.
#pragma once

#include "cdet_device.hpp"
#include "cdet_scalars.hpp"

namespace cdet {

constexpr double kUnitRoundoff = 1.1102230246251565e-16;  // 2^-53
constexpr int kBoundThreads = 256;
constexpr int kBoundGrid = 128;  // workgroups of the error-bound kernel (scratch slots)

__host__ __device__ inline size_t bound_doubles(int n) {
    const size_t sc = size_t(1) << n;
    const size_t n2 = static_cast<size_t>(n) * n;
    const size_t list = static_cast<size_t>(fpm_list_size(n));
    const size_t q = n >= 2 ? (size_t(1) << (n - 2)) : 1;
    return 2 * n2 + 4 * list + 2 * q + 2 * sc + 2 * sc + 4 * sc + 4 * sc + 4 * static_cast<size_t>(n) + 64 +
           2 * n2 * kBoundThreads;
}

struct BoundBufs {
    double *A, *EA, *L0v, *L0e, *L1v, *L1e, *invv, *inve, *pivv, *pive, *prv, *pre;
    double *pmUv, *pmUe, *pmDv, *pmDe, *av, *ae, *dpv, *dpe, *dU, *dD, *EdU, *EdD, *W, *EW;
    __device__ inline void carve(double* base, int n) {
        const size_t sc = size_t(1) << n;
        const size_t n2 = static_cast<size_t>(n) * n;
        const size_t list = static_cast<size_t>(fpm_list_size(n));
        const size_t q = n >= 2 ? (size_t(1) << (n - 2)) : 1;
        double* p = base;
        auto take = [&p](size_t k) {
            double* r = p;
            p += k;
            return r;
        };
        A = take(n2);
        EA = take(n2);
        L0v = take(list);
        L0e = take(list);
        L1v = take(list);
        L1e = take(list);
        invv = take(q);
        inve = take(q);
        pivv = take(sc);
        pive = take(sc);
        prv = take(sc);
        pre = take(sc);
        pmUv = take(sc);
        pmUe = take(sc);
        pmDv = take(sc);
        pmDe = take(sc);
        av = take(sc);
        ae = take(sc);
        dpv = take(sc);
        dpe = take(sc);
        dU = take(n);
        dD = take(n);
        EdU = take(n);
        EdD = take(n);
        take(64);
        W = take(n2 * kBoundThreads);
        EW = take(n2 * kBoundThreads);
    }
};

// Minors of A (values pmv, bounds pme) by the shifted Schur-complement tree;
// returns 1, leaving them undefined, when a pivot trips the fallback test.
__device__ inline int fpm_bound(const BoundBufs& b, int n, double tol, double thresh, double* pmv, double* pme,
                                double* red, int* flag) {
    const double u = kUnitRoundoff;
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const uint32_t sc = 1u << n;
    const int msize = n * n;
    double lmax = 0.0;
    for (int i = tid; i < msize; i += nt) lmax = fmax(lmax, fabs(b.A[i]));
    const double max_abs = block_max(lmax, red);
    const double p = !(thresh < 0.0) ? thresh : max_abs / static_cast<double>(msize);
    if (max_abs == 0.0 && p == 0.0) {
        for (uint32_t m = tid; m < sc; m += nt) {
            pmv[m] = m == 0 ? 1.0 : 0.0;
            pme[m] = 0.0;
        }
        __syncthreads();
        return 0;
    }
    const double fb_eps = (tol > 0.0 && max_abs > 0.0) ? tol * max_abs : 0.0;
    for (int i = tid; i < msize; i += nt) {
        b.L0v[i] = b.A[i];
        b.L0e[i] = b.EA[i];
    }
    for (uint32_t m = tid; m < sc; m += nt) {
        const double v = m == 0 ? 1.0 : (m == 1 ? p + b.A[0] : p);
        b.pivv[m] = v;
        b.pive[m] = m == 1 ? b.EA[0] + u * fabs(v) : 0.0;
    }
    if (tid == 0) *flag = 0;
    __syncthreads();
    double *ov = b.L0v, *oe = b.L0e, *nv = b.L1v, *ne = b.L1e;
    for (int l = 0; l < n - 1; ++l) {
        const int os = n - l;
        const int oms = os * os;
        const int ns = os - 1;
        const int nms = ns * ns;
        const int sets = 1 << l;
        for (int s = tid; s < sets; s += nt) {
            const double a00 = ov[s * oms];
            const double e00 = oe[s * oms];
            if (l >= 1) {
                const double v = p + a00;
                b.pivv[(1u << l) + s] = v;
                b.pive[(1u << l) + s] = e00 + u * fabs(v);
            }
            const double pd = a00 + p;
            if (fabs(pd) == 0.0 || (fb_eps > 0.0 && fabs(pd) <= fb_eps)) *flag = 1;
            const double inv = 1.0 / pd;
            b.invv[s] = inv;
            b.inve[s] = (e00 + u * fabs(pd)) * inv * inv + u * fabs(inv);
        }
        __syncthreads();
        if (*flag) return 1;
        const int items = sets * nms;
        for (int it = tid; it < items; it += nt) {
            const int s = it / nms;
            const int rem = it - s * nms;
            const int r = rem / ns;
            const int c = rem - r * ns;
            const int src = s * oms;
            const double x10 = ov[src + (r + 1) * os];
            const double e10 = oe[src + (r + 1) * os];
            const double f = x10 * b.invv[s];
            const double ef = fabs(b.invv[s]) * e10 + fabs(x10) * b.inve[s] + u * fabs(f);
            const double base = ov[src + (r + 1) * os + (c + 1)];
            const double eb = oe[src + (r + 1) * os + (c + 1)];
            const double x01 = ov[src + 1 + c];
            const double e01 = oe[src + 1 + c];
            nv[s * nms + rem] = base;
            ne[s * nms + rem] = eb;
            const double w = __builtin_fma(-x01, f, base);
            nv[(s + sets) * nms + rem] = w;
            ne[(s + sets) * nms + rem] = eb + fabs(x01) * ef + fabs(f) * e01 + u * fabs(w);
        }
        __syncthreads();
        double* t = ov;
        ov = nv;
        nv = t;
        t = oe;
        oe = ne;
        ne = t;
    }
    if (n >= 2) {
        const uint32_t top = 1u << (n - 1);
        for (uint32_t t = tid; t < top; t += nt) {
            const double v = p + ov[t];
            b.pivv[top + t] = v;
            b.pive[top + t] = oe[t] + u * fabs(v);
        }
    }
    __syncthreads();
    // prefix products over the bits of m from the lowest upward
    for (uint32_t m = tid; m < sc; m += nt) {
        double val = 1.0, e = 0.0;
        if (m != 0) {
            uint32_t pre = m & (~m + 1u);
            uint32_t rest = m ^ pre;
            val = b.pivv[pre];
            e = b.pive[pre];
            while (rest) {
                const uint32_t bit = rest & (~rest + 1u);
                rest ^= bit;
                pre |= bit;
                const double x = b.pivv[pre] * val;
                e = fabs(b.pivv[pre]) * e + fabs(val) * b.pive[pre] + u * fabs(x);
                val = x;
            }
        }
        pmv[m] = val;
        pme[m] = e;
    }
    __syncthreads();
    // removal of the uniform shift, highest bit first
    for (int l = n; l > 0; --l) {
        const uint32_t H = 1u << (l - 1);
        for (uint32_t g = tid; g < (sc >> 1); g += nt) {
            const uint32_t m = ((g >> (l - 1)) << l) | H | (g & (H - 1));
            const double x = __builtin_fma(-p, pmv[m - H], pmv[m]);
            pme[m] = pme[m] + p * pme[m - H] + u * fabs(x);
            pmv[m] = x;
        }
        __syncthreads();
    }
    return 0;
}

// Pivoted determinant of A[mask, mask] with its bound; the thread's k x k
// working matrices sit at w, ew with element (r, c) at [(r k + c) ws].
__device__ inline void det_bound(const double* A, const double* EA, int n, uint32_t mask, double* w, double* ew,
                                 int ws, double& det_out, double& e_out) {
    const double u = kUnitRoundoff;
    const int k = __builtin_popcount(mask);
    if (k == 0) {
        det_out = 1.0;
        e_out = 0.0;
        return;
    }
    int r = 0;
    for (int i = 0; i < n; ++i) {
        if (!(mask & (1u << i))) continue;
        int cc = 0;
        for (int j = 0; j < n; ++j) {
            if (!(mask & (1u << j))) continue;
            w[(r * k + cc) * ws] = A[i * n + j];
            ew[(r * k + cc) * ws] = EA[i * n + j];
            ++cc;
        }
        ++r;
    }
    double det = 1.0, edet = 0.0;
    int sign = 1;
    for (int col = 0; col < k; ++col) {
        int prow = col;
        double pabs = fabs(w[(col * k + col) * ws]);
        for (int rr = col + 1; rr < k; ++rr) {
            const double cand = fabs(w[(rr * k + col) * ws]);
            if (cand > pabs) {
                pabs = cand;
                prow = rr;
            }
        }
        if (pabs == 0.0) {  // exactly singular: no first-order bound from the elimination
            det_out = 0.0;
            e_out = __builtin_inf();
            return;
        }
        if (prow != col) {
            for (int c = col; c < k; ++c) {
                const double t = w[(col * k + c) * ws];
                w[(col * k + c) * ws] = w[(prow * k + c) * ws];
                w[(prow * k + c) * ws] = t;
                const double te = ew[(col * k + c) * ws];
                ew[(col * k + c) * ws] = ew[(prow * k + c) * ws];
                ew[(prow * k + c) * ws] = te;
            }
            sign = -sign;
        }
        const double piv = w[(col * k + col) * ws];
        const double epiv = ew[(col * k + col) * ws];
        const double nd = det * piv;
        edet = fabs(piv) * edet + fabs(det) * epiv + u * fabs(nd);
        det = nd;
        for (int rr = col + 1; rr < k; ++rr) {
            const double x = w[(rr * k + col) * ws];
            const double f = x / piv;
            const double ef = ew[(rr * k + col) * ws] / fabs(piv) + fabs(x) * epiv / (piv * piv) + u * fabs(f);
            for (int c = col + 1; c < k; ++c) {
                const double y = w[(col * k + c) * ws];
                const double z = __builtin_fma(-f, y, w[(rr * k + c) * ws]);
                ew[(rr * k + c) * ws] = ew[(rr * k + c) * ws] + fabs(f) * ew[(col * k + c) * ws] + fabs(y) * ef +
                                        u * fabs(z);
                w[(rr * k + c) * ws] = z;
            }
        }
    }
    det_out = sign < 0 ? -det : det;
    e_out = edet;
}

// Error bounds of the all-subset values of one class of each configuration:
// out (twisted) / out_ord (ordinary) at the configuration's output offset;
// the caller adds them as twisted + 2 ordinary.  e_up / e_dn carry the
// table-value error scale of every Green pair (same index as g_up / g_dn).
__global__ __launch_bounds__(kBoundThreads) void k_bound(KScalars ks, int n, int items,
                                                          const int32_t* __restrict__ cfg_list,
                                                          const int64_t* __restrict__ cfg_pbase,
                                                          const int64_t* __restrict__ cfg_obase,
                                                          const double* __restrict__ g_up,
                                                          const double* __restrict__ g_dn,
                                                          const double* __restrict__ e_up,
                                                          const double* __restrict__ e_dn, double* __restrict__ out,
                                                          double* __restrict__ out_ord,
                                                          const uint16_t* __restrict__ masks_by_pop,
                                                          const int* __restrict__ pop_off, double* __restrict__ gscratch,
                                                          size_t stride) {
    __shared__ double red[kBoundThreads / 64];
    __shared__ int s_flag;
    const double u = kUnitRoundoff;
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const uint32_t sc = 1u << n;
    const int n2 = n * n;
    BoundBufs b;
    b.carve(gscratch + static_cast<size_t>(blockIdx.x) * stride, n);
    for (int w = blockIdx.x; w < items; w += gridDim.x) {
        const int li = w >> 1;
        const int cls = w & 1;
        const int c = cfg_list[li];
        const bool same = ks.same_spin[cls] != 0;
        const int64_t pb = 2 * cfg_pbase[c] + static_cast<int64_t>(cls) * n2;
        for (int spin = 0; spin < (same ? 1 : 2); ++spin) {
            const double* g = (spin ? g_dn : g_up) + pb;
            const double* eg = (spin ? e_dn : e_up) + pb;
            const double shift = spin ? ks.ddn : ks.dup;
            for (int idx = tid; idx < n2; idx += nt) {
                const int i = idx / n;
                double v = g[idx];
                double e = eg[idx];
                if (i * n + i == idx) {
                    v += shift;
                    e += u * fabs(v);
                }
                const double a = v * ks.scale;
                b.A[idx] = a;
                b.EA[idx] = fabs(ks.scale) * e + u * fabs(a);
            }
            __syncthreads();
            double* pv = spin ? b.pmDv : b.pmUv;
            double* pe = spin ? b.pmDe : b.pmUe;
            if (fpm_bound(b, n, ks.tol, ks.shift, b.prv, b.pre, red, &s_flag)) {
                for (uint32_t m = tid; m < sc; m += nt) {
                    double d, e;
                    det_bound(b.A, b.EA, n, m, b.W + tid, b.EW + tid, nt, d, e);
                    pv[m] = d;
                    pe[m] = e;
                }
            } else {
                for (uint32_t m = tid; m < sc; m += nt) {
                    pv[m] = b.prv[m];
                    pe[m] = b.pre[m];
                }
            }
            __syncthreads();
        }
        const double* pUv = b.pmUv;
        const double* pUe = b.pmUe;
        const double* pDv = same ? b.pmUv : b.pmDv;
        const double* pDe = same ? b.pmUe : b.pmDe;
        for (int i = tid; i < n; i += nt) {
            const double du = g_up[pb + i * n + i] + ks.dup;
            const double dd = g_dn[pb + i * n + i] + ks.ddn;
            b.dU[i] = du;
            b.dD[i] = dd;
            b.EdU[i] = e_up[pb + i * n + i] + u * fabs(du);
            b.EdD[i] = e_dn[pb + i * n + i] + u * fabs(dd);
        }
        for (uint32_t m = tid; m < sc; m += nt) {
            if (m == 0) {
                b.av[0] = 0.0;
                b.ae[0] = 0.0;
            } else {
                const double a = pUv[m] * pDv[m];
                b.av[m] = a;
                b.ae[m] = fabs(pUv[m]) * pDe[m] + fabs(pDv[m]) * pUe[m] + u * fabs(a);
            }
            const bool single = m != 0 && (m & (m - 1)) == 0;
            b.dpv[m] = single ? b.av[m] : 0.0;
            b.dpe[m] = single ? b.ae[m] : 0.0;
        }
        __syncthreads();
        // all-subset recursion, one mask per lane, popcount layer by layer
        for (int bits = 2; bits <= n; ++bits) {
            const int beg = pop_off[bits];
            const int end = pop_off[bits + 1];
            const int terms = (1 << (bits - 1)) - 1;
            for (int it = beg + tid; it < end; it += nt) {
                const uint32_t mask = masks_by_pop[it];
                const uint32_t anchor = mask & (~mask + 1u);
                const uint32_t rest = mask ^ anchor;
                double acc = b.av[mask];
                double eacc = b.ae[mask];
                double rhs = fabs(acc);
                uint32_t sub_s = (rest - 1u) & rest;
                for (int t = 0; t < terms; ++t) {
                    const uint32_t sub = anchor | sub_s;
                    const double dv = b.dpv[sub];
                    const double ar = b.av[mask ^ sub];
                    const double term = dv * ar;
                    acc = acc - term;
                    eacc += fabs(dv) * b.ae[mask ^ sub] + fabs(ar) * b.dpe[sub] + u * fabs(term) + u * fabs(acc);
                    rhs += fabs(term);
                    sub_s = (sub_s - 1u) & rest;
                }
                double val = acc;
                double e = eacc;
                if (ks.filter) {
                    const double window = ks.thr * (rhs + eacc) + eacc;
                    if (fabs(acc) < ks.thr * rhs) val = 0.0;
                    if (fabs(acc) <= window) e = window;
                }
                b.dpv[mask] = val;
                b.dpe[mask] = e;
            }
            __syncthreads();
        }
        // assembly: bounds of the class values
        double* dst = (cls == 0 ? out : out_ord) + cfg_obase[c];
        for (uint32_t m = tid; m < sc; m += nt) {
            double e;
            if (m == 0) {
                e = 0.0;
            } else if ((m & (m - 1)) == 0) {
                const int i = __builtin_ctz(m);
                const double prod = b.dU[i] * b.dD[i];
                const double diff = prod - ks.counter;
                const double v = ks.c1[cls] * diff;
                e = fabs(ks.c1[cls]) * (fabs(b.dD[i]) * b.EdU[i] + fabs(b.dU[i]) * b.EdD[i] + u * fabs(prod) +
                                        u * fabs(diff)) +
                    u * fabs(v);
            } else {
                const double pf = ks.pref[cls][__builtin_popcount(m)];
                e = fabs(pf) * b.dpe[m] + u * fabs(pf * b.dpv[m]);
            }
            dst[m] = e;
        }
        __syncthreads();
    }
}

}  // namespace cdet

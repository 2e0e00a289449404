// cdet_cdet.hpp -- connected-determinant subset recursions,
// workgroup-cooperative, FP64 in the reference's operation order.
#pragma once

#include "cdet_device.hpp"

namespace cdet {

// One mask of the subset recursion: acc = a[mask] - sum over the
// anchor-containing proper subsets `anchor | s` (s over the proper subsets of
// rest, descending) of dp[sub] * a[mask ^ sub], with the running |term| sum
// in rhs_out, exactly the reference's sequence
// (include/cdet/integrand_replica_randomly_sampled.h:326-343).  Operands are
// read eight terms at a time, one chunk ahead of the dependent subtractions,
// so that LDS latency overlaps the serial chain; the arithmetic order is
// unchanged.
__device__ inline double connected_chain(const double* a, const double* dp, uint32_t mask, uint32_t anchor,
                                         uint32_t rest, int terms, double thr, bool filter, double& rhs_out) {
    double acc = a[mask];
    double rhs = fabs(acc);
    uint32_t s = (rest - 1u) & rest;
    int done = 0;
    if (terms >= 8) {
        double dv[8], av[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint32_t sub = anchor | s;
            dv[u] = dp[sub];
            av[u] = a[mask ^ sub];
            s = (s - 1u) & rest;
        }
        for (; done + 16 <= terms; done += 8) {
            double dn[8], an[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t sub = anchor | s;
                dn[u] = dp[sub];
                an[u] = a[mask ^ sub];
                s = (s - 1u) & rest;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const double term = dv[u] * av[u];
                acc = acc - term;
                rhs += fabs(term);
                dv[u] = dn[u];
                av[u] = an[u];
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const double term = dv[u] * av[u];
            acc = acc - term;
            rhs += fabs(term);
        }
        done += 8;
    }
    const int tail = terms - done;  // < 8
    if (tail > 0) {
        double dv[7], av[7];
#pragma unroll
        for (int u = 0; u < 7; ++u) {
            if (u < tail) {
                const uint32_t sub = anchor | s;
                dv[u] = dp[sub];
                av[u] = a[mask ^ sub];
                s = (s - 1u) & rest;
            }
        }
#pragma unroll
        for (int u = 0; u < 7; ++u) {
            if (u < tail) {
                const double term = dv[u] * av[u];
                acc = acc - term;
                rhs += fabs(term);
            }
        }
    }
    if (filter && fabs(acc) < thr * rhs) acc = 0.0;
    rhs_out = rhs;
    return acc;
}

// All-subset recursion (include/cdet/integrand_replica_randomly_sampled.h:292-347).
// On entry a[mask] = pm_up * pm_down (a[0] = 0); on exit dp[mask] holds the
// connected value.  masks_by_pop lists every mask ordered by popcount with
// pop_off[k] the first index of popcount k; masks of one popcount are
// independent and run in parallel.
// With rhs non-null, rhs[mask] receives the sum of |a[mask]| and |terms| of
// the mask's chain (the scale against which the round-off filter judges the
// connected value, :330-341).
__device__ inline void ztransform_all_block(const double* a, double* dp, int n, double thr, bool filter,
                                            const uint16_t* masks_by_pop, const int* pop_off, double* rhs = nullptr) {
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const uint32_t sc = 1u << n;
    for (uint32_t m = tid; m < sc; m += nt) dp[m] = (m != 0 && (m & (m - 1)) == 0) ? a[m] : 0.0;
    __syncthreads();
    for (int bits = 2; bits <= n; ++bits) {
        const int beg = pop_off[bits];
        const int end = pop_off[bits + 1];
        const int terms = (1 << (bits - 1)) - 1;
        for (int it = beg + tid; it < end; it += nt) {
            const uint32_t mask = masks_by_pop[it];
            const uint32_t anchor = mask & (~mask + 1u);
            double r;
            dp[mask] = connected_chain(a, dp, mask, anchor, mask ^ anchor, terms, thr, filter, r);
            if (rhs) rhs[mask] = r;
        }
        __syncthreads();
    }
}

// Single-anchor recursion (include/cdet/integrand_replica_randomly_sampled.h:226-290):
// only masks containing bit 0 recurse; returns dp[full] (valid on all threads).
__device__ inline double ztransform_single_block(const double* a, double* dp, int n, double thr, bool filter,
                                                 const uint16_t* masks_by_pop, const int* pop_off) {
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const uint32_t sc = 1u << n;
    for (uint32_t m = tid; m < sc; m += nt) dp[m] = (m != 0 && (m & (m - 1)) == 0) ? a[m] : 0.0;
    __syncthreads();
    for (int bits = 2; bits <= n; ++bits) {
        const int beg = pop_off[bits];
        const int end = pop_off[bits + 1];
        const int terms = (1 << (bits - 1)) - 1;
        for (int it = beg + tid; it < end; it += nt) {
            const uint32_t mask = masks_by_pop[it];
            double r;
            dp[mask] = (mask & 1u) ? connected_chain(a, dp, mask, 1u, mask ^ 1u, terms, thr, filter, r) : a[mask];
        }
        __syncthreads();
    }
    const double r = dp[sc - 1];
    __syncthreads();
    return r;
}

}  // namespace cdet

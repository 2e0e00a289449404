// cdet_cdet.hpp -- connected-determinant subset recursions and output
// assembly, workgroup-cooperative, FP64 in the reference's operation order.
#pragma once

#include "cdet_device.hpp"

namespace cdet {

// All-subset recursion (include/cdet/integrand_replica_randomly_sampled.h:292-347).
// On entry a[mask] = pm_up * pm_down (a[0] = 0); on exit dp[mask] holds the
// connected value.  masks_by_pop lists every mask ordered by popcount with
// pop_off[k] the first index of popcount k.  Each mask is one sequential
// chain of subtractions over its anchor-containing proper subsets in
// descending order, exactly as the reference iterates them; masks of equal
// popcount are independent and run in parallel.
__device__ inline void ztransform_all_block(const double* a, double* dp, int n, double thr, bool filter,
                                            const uint16_t* masks_by_pop, const int* pop_off) {
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const uint32_t sc = 1u << n;
    for (uint32_t m = tid; m < sc; m += nt) dp[m] = (m != 0 && (m & (m - 1)) == 0) ? a[m] : 0.0;
    __syncthreads();
    for (int bits = 2; bits <= n; ++bits) {
        const int beg = pop_off[bits];
        const int end = pop_off[bits + 1];
        for (int it = beg + tid; it < end; it += nt) {
            const uint32_t mask = masks_by_pop[it];
            const uint32_t anchor = mask & (~mask + 1u);
            const uint32_t rest = mask ^ anchor;
            double acc = a[mask];
            double rhs = fabs(a[mask]);
            // subsets containing the anchor, descending: anchor | s, s over
            // proper subsets of rest in descending order.
            for (uint32_t s = (rest - 1u) & rest;; s = (s - 1u) & rest) {
                const uint32_t sub = anchor | s;
                const double term = dp[sub] * a[mask ^ sub];
                acc = acc - term;
                rhs += fabs(term);
                if (s == 0u) break;
            }
            if (filter && fabs(acc) < thr * rhs) acc = 0.0;
            dp[mask] = acc;
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
        for (int it = beg + tid; it < end; it += nt) {
            const uint32_t mask = masks_by_pop[it];
            double acc = a[mask];
            if (mask & 1u) {
                double rhs = fabs(a[mask]);
                const uint32_t rest = mask ^ 1u;
                for (uint32_t s = (rest - 1u) & rest;; s = (s - 1u) & rest) {
                    const uint32_t sub = 1u | s;
                    const double term = dp[sub] * a[mask ^ sub];
                    acc = acc - term;
                    rhs += fabs(term);
                    if (s == 0u) break;
                }
                if (filter && fabs(acc) < thr * rhs) acc = 0.0;
            }
            dp[mask] = acc;
        }
        __syncthreads();
    }
    const double r = dp[sc - 1];
    __syncthreads();
    return r;
}

}  // namespace cdet

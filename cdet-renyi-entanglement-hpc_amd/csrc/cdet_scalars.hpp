// cdet_scalars.hpp -- scalar inputs and table pointers as the kernels see them.
#pragma once

#include "cdet_device.hpp"
#include "../../include/cdet_hip.h"

namespace cdet {

struct KScalars {
    int N1, N2, nu, nv;
    int64_t SP;              // sitepairs per table kind, 4 (N1 N2)^2
    double beta, U, branch_eps, tol, shift, thr, scale, dup, ddn, counter;
    int filter;
    double c1[2];                      // ((-rc) * U) * beta, class 0 (rc=2) / class 1 (rc=1)
    double pref[2][kMaxOrder + 1];     // all-subset prefactor (long double, cast)
    double pref_hi[2][kMaxOrder + 1];  // single-order prefactor, double-double
    double pref_lo[2][kMaxOrder + 1];
    int same_spin[2];                  // class up/down tables and shifts identical
    int group16_R[kMaxOrder + 1];      // per order: chain length from which <= 16-mask layers group 16 lanes
    int group64_R[kMaxOrder + 1];      // per order: chain length from which a single chain takes a wave
    int tree;                          // recursion sums in tree order (tree_layer), not the reference's
    uint64_t tree_bits[kMaxOrder + 1];  // per order n: bits 3k..3k+2 = log2 lanes per mask of layer k
    // Order n's layer shapes when the sums run in tree order, else 0 (a value,
    // not a pointer into the argument block, which would put the block in
    // scratch memory).
    __device__ inline uint64_t tree_row(int n) const { return tree ? tree_bits[n] : 0; }
    const double* coeff[4][2];         // [kind][branch 0 = upper, 1 = lower]
    const double* corner[4][2];
    const double* steps[4][2];         // coeff in matrix-core step order (FP64 MFMA mode)
    const float* split_hi[4][2];       // FP32 hi/lo split of coeff (split-FP32 mode)
    const float* split_lo[4][2];
    const double* abssum[4][2];        // per coefficient block, sum of |c_ab| (error-bound kind)
    const double* rowsum[4];           // lower branch, per block row a: sum over b of c_ab (FP64 matrix-core mode)

    // Coefficient block of a query key (class, branch, sitepair).
    __device__ inline void key_tables(int32_t key, const double*& cu, const double*& cd, bool& same) const {
        const int64_t sp = key % SP;
        const int rest = static_cast<int>(key / SP);
        const int br = rest & 1;
        const int cls = rest >> 1;
        const int ku = cls == 0 ? CDET_TABLE_TWISTED_UP : CDET_TABLE_ORDINARY_UP;
        const int64_t base = sp * static_cast<int64_t>(nu) * nv;
        cu = coeff[ku][br] + base;
        cd = coeff[ku + 1][br] + base;
        same = same_spin[cls] != 0;
    }
    // The same block in matrix-core step order.
    __device__ inline void key_steps(int32_t key, const double*& cu, const double*& cd, bool& same) const {
        const int64_t sp = key % SP;
        const int rest = static_cast<int>(key / SP);
        const int br = rest & 1;
        const int cls = rest >> 1;
        const int ku = cls == 0 ? CDET_TABLE_TWISTED_UP : CDET_TABLE_ORDINARY_UP;
        const int64_t base = sp * static_cast<int64_t>(nu) * nv;
        cu = steps[ku][br] + base;
        cd = steps[ku + 1][br] + base;
        same = same_spin[cls] != 0;
    }
    __device__ inline void key_split(int32_t key, const float*& hu, const float*& lu, const float*& hd,
                                     const float*& ld) const {
        const int64_t sp = key % SP;
        const int rest = static_cast<int>(key / SP);
        const int br = rest & 1;
        const int cls = rest >> 1;
        const int ku = cls == 0 ? CDET_TABLE_TWISTED_UP : CDET_TABLE_ORDINARY_UP;
        const int64_t base = sp * static_cast<int64_t>(nu) * nv;
        hu = split_hi[ku][br] + base;
        lu = split_lo[ku][br] + base;
        hd = split_hi[ku + 1][br] + base;
        ld = split_lo[ku + 1][br] + base;
    }
};

}  // namespace cdet

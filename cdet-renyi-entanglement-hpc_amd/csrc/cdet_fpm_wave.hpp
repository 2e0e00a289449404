// cdet_fpm_wave.hpp -- principal minors of one n x n matrix (n <= 10) on one
// wavefront, with no workgroup barrier.
//
// The shifted Schur-complement tree of the reference's mat2pm
// (include/cdet/fast_principle_minors.h:120-246) has n - 1 levels; level l
// holds 2^l matrices of size n - l, and matrix s of level l has the children s
// ("without", its trailing copy) and s + 2^l ("with", its Schur complement).
// Every element's value depends only on its ancestors, so the tree may be
// traversed in any order with the same arithmetic, value for value:
//   * levels 0 .. L-1 (L = min(6, n - 1)) run breadth-first in LDS, the 64
//     lanes sharing each level's elements and synchronising within the wave;
//   * at level L each lane s < 2^L takes matrix s and runs its whole subtree
//     (levels L .. n-1, matrices s + 2^L k) in registers: lane s ends up with
//     the pivots of every mask whose low L bits are s;
//   * the prefix products final[m] = out[m] * final[m - topbit(m)] and the
//     removal of the uniform shift (bit by bit, highest first:
//     b[m] = fma(-p, b[m - 2^b], b[m]) for masks with bit b set) then run on
//     those 2^(n-L) values per lane, the low bits across lanes by wave
//     shuffles.
// The operations on each element are those of fpm_block (and of the
// reference): pd = a00 + p, inv = 1 / pd, f = x(r+1, 0) * inv,
// with = fma(-x(0, c+1), f, x(r+1, c+1)), out = p + a00, the same pivot test
// against TOL_FPM * max|A|, and the same product and shift-removal sequence.
#pragma once

#include "cdet_device.hpp"
#include "cdet_scalars.hpp"

namespace cdet {

constexpr int kWaveFpmMaxOrder = 10;
constexpr int kWaveFpmL = 6;  // level at which lanes take one matrix each (2^6 = 64 lanes)

__host__ __device__ constexpr int fpm_wave_L(int n) { return n - 1 < kWaveFpmL ? n - 1 : kWaveFpmL; }

// LDS of the wave FPM: two level lists (even / odd levels up to L), the
// inverse pivots of a level (<= 2^(L-1)) and the pivots of masks < 2^L.
// fpm_wave_e0: doubles of the even-level list; fpm_wave_e1: the odd-level
// list with the 32 inverse pivots and 64 low-mask pivots behind it.
__host__ __device__ constexpr int fpm_wave_e0(int n) {
    int e = 0;
    for (int l = 0; l <= fpm_wave_L(n); l += 2) {
        const int sz = (1 << l) * (n - l) * (n - l);
        e = sz > e ? sz : e;
    }
    return e;
}
__host__ __device__ constexpr int fpm_wave_e1(int n) {
    int e = 0;
    for (int l = 1; l <= fpm_wave_L(n); l += 2) {
        const int sz = (1 << l) * (n - l) * (n - l);
        e = sz > e ? sz : e;
    }
    return e + 32 + 64;
}
__host__ __device__ constexpr int fpm_wave_lds_doubles(int n) { return fpm_wave_e0(n) + fpm_wave_e1(n); }

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__host__ __device__ constexpr int fpm_sub_size(int M) {
    int best = 1;
    for (int j = 0; j < M; ++j) {
        const int sz = (1 << j) * (M - j) * (M - j);
        best = sz > best ? sz : best;
    }
    return best;
}

// Subtree of one level-L matrix X (M x M, row-major) in registers: piv[h] for
// h in [1, 2^M) is out[2^L h + s]: h = 2^j + k is matrix k of sub-level j
// (global level L + j).  Sets bad when a pivot trips the fallback test.
template <int M>
__device__ __forceinline__ void fpm_lane_subtree(const double (&X)[M * M], double p, double fb_eps,
                                                 double (&piv)[1 << M], bool& bad) {
    constexpr int S = fpm_sub_size(M);
    double cur[S], nxt[S];
#pragma unroll
    for (int i = 0; i < M * M; ++i) cur[i] = X[i];
#pragma unroll
    for (int j = 0; j < M; ++j) {
        const int m = M - j;
        const int cnt = 1 << j;
        if (j == M - 1) {
            // final 1 x 1 matrices: pivot value only, no test (fpm_block's
            // out[top + t] += oldl[t])
#pragma unroll
            for (int k = 0; k < cnt; ++k) piv[cnt + k] = p + cur[k];
        } else {
            const int ns = m - 1;
#pragma unroll
            for (int k = 0; k < cnt; ++k) {
                const double a00 = cur[k * m * m];
                piv[cnt + k] = p + a00;
                const double pd = a00 + p;
                const double pda = fabs(pd);
                if (pda == 0.0 || (fb_eps > 0.0 && pda <= fb_eps)) bad = true;
                const double inv = 1.0 / pd;
#pragma unroll
                for (int r = 0; r < ns; ++r) {
                    const double f = cur[k * m * m + (r + 1) * m] * inv;
#pragma unroll
                    for (int c = 0; c < ns; ++c) {
                        const double base = cur[k * m * m + (r + 1) * m + (c + 1)];
                        nxt[k * ns * ns + r * ns + c] = base;
                        nxt[(k + cnt) * ns * ns + r * ns + c] = __builtin_fma(-cur[k * m * m + 1 + c], f, base);
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < S; ++i) cur[i] = nxt[i];
        }
    }
}

// Outcome of fpm_wave_minors.
constexpr int kFpmWaveOk = 0;
constexpr int kFpmWaveFallback = 1;  // a pivot tripped the reference's fallback test

// Principal minors of one spin matrix on the calling wavefront.  g holds the
// n x n Green values (row-major) and shift the spin's diagonal shift; the
// vertex matrix is (g + shift I) scale
// (include/cdet/integrand_replica_randomly_sampled.h:751-761).  list0
// (fpm_wave_e0(n) doubles) and list1 (fpm_wave_e1(n) doubles) are LDS used by
// this wave alone; the wave's last LDS access precedes the return.  On
// kFpmWaveOk, lane s < 2^L holds b[h] = det of the principal submatrix of mask
// 2^L h + s (h < 2^M, M = n - L); an all-zero matrix gives 1 for mask 0 and 0
// elsewhere, as mat2pm does.  kFpmWaveFallback is wave-uniform; b is then
// undefined.  Nothing is written outside scr.
template <int M>
__device__ __forceinline__ int fpm_wave_minors(const KScalars& ks, int n, const double* __restrict__ g, double shift,
                                               double* list0, double* list1, int lane, double (&b)[1 << M]) {
    const int L = fpm_wave_L(n);  // = n - M
    const int T = 1 << L;
    const int n2 = n * n;
    double* inv = list1 + (fpm_wave_e1(n) - 96);
    double* outlow = inv + 32;
    double lmax = 0.0;
    for (int idx = lane; idx < n2; idx += 64) {
        const int i = idx / n;
        double v = g[idx];
        if (i * n + i == idx) v += shift;
        const double a = v * ks.scale;
        list0[idx] = a;
        const double e = fabs(a);
        lmax = e > lmax ? e : lmax;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(lmax, off);
        lmax = o > lmax ? o : lmax;
    }
    const double max_abs = lmax;
    const bool explicit_shift = !(ks.shift < 0.0);
    const double p = explicit_shift ? ks.shift : max_abs / static_cast<double>(n2);
    if (max_abs == 0.0 && p == 0.0) {
#pragma unroll
        for (int h = 0; h < (1 << M); ++h) b[h] = (h == 0 && lane == 0) ? 1.0 : 0.0;
        return kFpmWaveOk;
    }
    const double fb_eps = (ks.tol > 0.0 && max_abs > 0.0) ? ks.tol * max_abs : 0.0;
    wave_sync_lds();
    if (lane < T) outlow[lane] = (lane == 0) ? 1.0 : (lane == 1 ? p + list0[0] : p);
    bool bad = false;
    // levels 0 .. L-1, breadth-first in LDS
    double* oldl = list0;
    double* newl = list1;
    for (int l = 0; l < L; ++l) {
        const int os = n - l;
        const int oms = os * os;
        const int ns = os - 1;
        const int nms = ns * ns;
        const int sets = 1 << l;
        if (lane < sets) {
            const double a00 = oldl[lane * oms];
            if (l >= 1) outlow[(1 << l) + lane] = p + a00;
            const double pd = a00 + p;
            const double pda = fabs(pd);
            if (pda == 0.0 || (fb_eps > 0.0 && pda <= fb_eps)) bad = true;
            inv[lane] = 1.0 / pd;
        }
        wave_sync_lds();
        const int nitems = sets * nms;
        // the last level (the lanes' level-L matrices) is stored element-major,
        // element e of matrix s at e 2^L + s, so that each lane's matrix load
        // below is one contiguous wave read per element (matrix-major, the
        // lanes' loads would be 8 M^2 bytes apart and hit the same LDS banks)
        const bool last = l == L - 1;
        // item it = s nms + rem, rem = r ns + cc: divided once for the lane's
        // first item, then stepped by the wave stride 64 = ds nms + drem,
        // drem = dr ns + dcc, with carries
        const int ds = 64 / nms;
        const int drem = 64 - ds * nms;
        const int dr = drem / ns;
        const int dcc = drem - dr * ns;
        int s = lane / nms;
        int rem = lane - s * nms;
        int r = rem / ns;
        int cc = rem - r * ns;
        for (int it = lane; it < nitems; it += 64) {
            const int src = s * oms;
            const double f = oldl[src + (r + 1) * os] * inv[s];
            const double base = oldl[src + (r + 1) * os + (cc + 1)];
            const double with = __builtin_fma(-oldl[src + 1 + cc], f, base);
            if (last) {
                newl[rem * (2 * sets) + s] = base;
                newl[rem * (2 * sets) + s + sets] = with;
            } else {
                newl[s * nms + rem] = base;
                newl[(s + sets) * nms + rem] = with;
            }
            cc += dcc;
            const int c2 = cc >= ns ? 1 : 0;
            cc -= c2 * ns;
            r += dr + c2;
            const int c1 = r >= ns ? 1 : 0;
            r -= c1 * ns;
            rem += drem;
            rem -= c1 * nms;
            s += ds + c1;
        }
        wave_sync_lds();
        double* t = oldl;
        oldl = newl;
        newl = t;
    }
    // levels L .. n-1: lane s < T takes level-L matrix s (M x M)
    double piv[1 << M];
    double X[M * M];
    const int s = lane < T ? lane : 0;
#pragma unroll
    for (int i = 0; i < M * M; ++i) X[i] = oldl[i * T + s];
    fpm_lane_subtree<M>(X, p, fb_eps, piv, bad);
    if (lane >= T) bad = false;
    // running product over the low bits of s (final[s], masks < 2^L)
    double fin = 1.0;
    if (s != 0) {
        uint32_t pre = static_cast<uint32_t>(s) & (~static_cast<uint32_t>(s) + 1u);
        uint32_t rest = static_cast<uint32_t>(s) ^ pre;
        fin = outlow[pre];
        while (rest) {
            const uint32_t bb = rest & (~rest + 1u);
            rest ^= bb;
            pre |= bb;
            fin = outlow[pre] * fin;
        }
    }
    if (__any(bad)) return kFpmWaveFallback;  // the reference switches to pivoted determinants
    // prefix products over the high bits: b[h] = final[2^L h + s]
    b[0] = fin;
#pragma unroll
    for (int h = 1; h < (1 << M); ++h) {
        int tb = 1;
        while ((tb << 1) <= h) tb <<= 1;
        b[h] = piv[h] * b[h ^ tb];
    }
    // shift removal, highest bit first: bits L .. n-1 within the lane, then
    // bits L-1 .. 0 across lanes
#pragma unroll
    for (int hb = M - 1; hb >= 0; --hb) {
#pragma unroll
        for (int h = 0; h < (1 << M); ++h)
            if ((h >> hb) & 1) b[h] = __builtin_fma(-p, b[h ^ (1 << hb)], b[h]);
    }
    for (int bit = L - 1; bit >= 0; --bit) {
        const int sb = 1 << bit;
        const bool set = (lane & sb) != 0;
#pragma unroll
        for (int h = 0; h < (1 << M); ++h) {
            const double o = __shfl_xor(b[h], sb);
            if (set) b[h] = __builtin_fma(-p, o, b[h]);
        }
    }
    return kFpmWaveOk;
}

// Minors of one spin matrix (order n, n - fpm_wave_L(n) = M) by the calling
// wave into dst[mask] (the mask's low L bits across lanes), or the matrix
// queued as fb_entry for the pivoted-determinant fallback.
template <int M>
__device__ __forceinline__ void fpm_wave_matrix(const KScalars& ks, int n, const double* g, double shift, double* lds,
                                                int lane, double* dst, int32_t* fb_list, int32_t* fb_count,
                                                int32_t fb_entry) {
    double b[1 << M];
    if (fpm_wave_minors<M>(ks, n, g, shift, lds, lds + fpm_wave_e0(n), lane, b) == kFpmWaveFallback) {
        if (lane == 0) {
            const int slot = atomicAdd(fb_count, 1);
            fb_list[slot] = fb_entry;
        }
        return;
    }
    const int T = 1 << fpm_wave_L(n);
    if (lane < T) {
#pragma unroll
        for (int h = 0; h < (1 << M); ++h) dst[static_cast<size_t>(h) * T + lane] = b[h];
    }
}

// One spin matrix per 64-thread workgroup.  Work item w as k_fpm: 2 (li,
// class) pairs x 2 spins, or x 1 when both classes have identical spin tables.
// Writes pm[(4 li + 2 class + spin) 2^pm_log + mask], or queues the matrix for
// the pivoted-determinant fallback exactly as k_fpm does.  With cfg_order
// non-null the configurations are of several orders: configuration c has
// order cfg_order[c] <= n, and n only sized the LDS.  M = 0 takes
// n - fpm_wave_L(n) per configuration (orders 2 .. kWaveFpmMaxOrder in one
// launch); otherwise every order of the launch has n - fpm_wave_L(n) = M.
template <int M>
__global__ __launch_bounds__(64) void k_fpm_wave(KScalars ks, int n, int items, const int32_t* __restrict__ cfg_list,
                                                 const int64_t* __restrict__ cfg_pbase,
                                                 const double* __restrict__ g_up, const double* __restrict__ g_dn,
                                                 double* __restrict__ pm, int32_t* __restrict__ fb_list,
                                                 int32_t* __restrict__ fb_count, const int32_t* __restrict__ cfg_order,
                                                 int pm_log) {
    extern __shared__ double lds_dyn[];
    const int lane = threadIdx.x;
    const int w = blockIdx.x;
    if (w >= items) return;
    const bool both_same = ks.same_spin[0] && ks.same_spin[1];
    const int li = both_same ? (w >> 1) : (w >> 2);
    const int cls = both_same ? (w & 1) : ((w >> 1) & 1);
    const int spin = both_same ? 0 : (w & 1);
    const bool same = ks.same_spin[cls] != 0;
    if (spin == 1 && same) return;  // the up minors serve both spins
    const int c = cfg_list[li];
    if (cfg_order) n = cfg_order[c];
    const double* g = (spin ? g_dn : g_up) + 2 * cfg_pbase[c] + static_cast<int64_t>(cls) * n * n;
    const double shift = spin ? ks.ddn : ks.dup;
    double* dst = pm + (static_cast<size_t>(4 * li + 2 * cls + spin) << pm_log);
    const int32_t entry = (li << 3) | (cls << 2) | (same ? 3 : (1 << spin));
    if constexpr (M == 0) {
        switch (n - fpm_wave_L(n)) {
            case 1:
                fpm_wave_matrix<1>(ks, n, g, shift, lds_dyn, lane, dst, fb_list, fb_count, entry);
                break;
            case 2:
                fpm_wave_matrix<2>(ks, n, g, shift, lds_dyn, lane, dst, fb_list, fb_count, entry);
                break;
            case 3:
                fpm_wave_matrix<3>(ks, n, g, shift, lds_dyn, lane, dst, fb_list, fb_count, entry);
                break;
            default:
                fpm_wave_matrix<4>(ks, n, g, shift, lds_dyn, lane, dst, fb_list, fb_count, entry);
                break;
        }
    } else {
        fpm_wave_matrix<M>(ks, n, g, shift, lds_dyn, lane, dst, fb_list, fb_count, entry);
    }
}

}  // namespace cdet

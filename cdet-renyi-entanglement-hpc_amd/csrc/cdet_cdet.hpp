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
// unchanged.  Since s is a subset of rest and the anchor bit is outside it,
// dp[anchor | s] = dp[anchor + s] and a[mask ^ (anchor | s)] = a[rest - s]:
// both operands are one add from a per-mask base, with s kept as a byte
// offset (8 s) and stepped as (8 s - 8) & 8 rest.
__device__ __forceinline__ double connected_chain(const double* a, const double* dp, uint32_t mask, uint32_t anchor,
                                         uint32_t rest, int terms, double thr, bool filter, double& rhs_out) {
    double acc = a[mask];
    double rhs = fabs(acc);
    const char* dpb = reinterpret_cast<const char*>(dp + anchor);
    const char* ab = reinterpret_cast<const char*>(a + rest);
    const uint32_t rest8 = rest << 3;
    uint32_t s8 = ((rest - 1u) & rest) << 3;
    auto dp_at = [dpb](uint32_t o) { return *reinterpret_cast<const double*>(dpb + o); };
    auto a_at = [ab](uint32_t o) { return *reinterpret_cast<const double*>(ab - o); };
    int done = 0;
    if (terms >= 8) {
        double dv[8], av[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            dv[u] = dp_at(s8);
            av[u] = a_at(s8);
            s8 = (s8 - 8u) & rest8;
        }
        for (; done + 16 <= terms; done += 8) {
            double dn[8], an[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                dn[u] = dp_at(s8);
                an[u] = a_at(s8);
                s8 = (s8 - 8u) & rest8;
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
                dv[u] = dp_at(s8);
                av[u] = a_at(s8);
                s8 = (s8 - 8u) & rest8;
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

constexpr int kFoldSteps = 2;  // steps per exchange of the grouped chains
constexpr int kFoldWaves = 4;  // waves that can run grouped chains (fold buffer: kFoldWaves x kFoldSteps x 64)

// Lanes per mask for a popcount layer of M masks with chains of R - 1 terms
// (0: one lane per mask).  A single mask with a long chain takes a whole wave
// (64 lanes); up to four masks per wave, on at most kFoldWaves waves, take
// 16 lanes each.  Per 64 terms a group of G lanes issues about 10 + 2 G wave
// instructions (the products, then the leader's fold), against 9 per term on
// one lane, so grouping shortens the chain about 3.5x (G = 16) with no more
// instructions in total when a one-lane-per-mask layer would leave most lanes
// of its wave idle.
// g16 / g64: the chain lengths from which the 16-lane and whole-wave forms
// are taken (64 and 128 by default; tunable per order).
__device__ inline int layer_group(int M, int R, int waves, int g16 = 64, int g64 = 128) {
    const int wf = waves < kFoldWaves ? waves : kFoldWaves;
    if (M == 1 && R >= g64) return 64;
    if (M <= 4 * wf && R >= g16) return 16;
    return 0;
}

// One popcount layer as grouped chains of G lanes per mask.  The chain's terms
// are enumerated as x = G X + j (descending x is descending X, then
// descending j): lane j of the group holds s_lo = the j-th subset of the
// lowest log2(G) bits of the rest and steps the high part s_hi = (s_hi - 1) &
// rest_hi.  Every kFoldSteps steps the lanes' products go to tb[] and the
// group's lane 0 folds them in descending order, so the arithmetic sequence is
// the reference's.  The one excluded subset (s = rest, the mask itself) is
// the first term of the first step and enters as 0, which leaves acc and rhs
// unchanged.  Waves without a mask return at once; the caller synchronises.
// One round of group_chains: the wave's kPerWave groups take masks base ..
// base + kPerWave - 1 (those below end); wtb is the wave's fold buffer.
template <int G>
__device__ __forceinline__ void group_chain_round(const double* a, double* dp, int base, int end, int R, double thr,
                                                  bool filter, const uint16_t* masks_by_pop, double* rhs, bool single,
                                                  double* wtb) {
    constexpr int kLoBits = G == 64 ? 6 : (G == 32 ? 5 : (G == 16 ? 4 : (G == 8 ? 3 : (G == 4 ? 2 : 1))));
    const int lane = threadIdx.x & 63;
    const int gi = lane / G;
    const int j = lane - gi * G;
    const int it = base + gi;
    const bool have = it < end;
    const uint32_t mask = have ? masks_by_pop[it] : 1u;
    const bool rec = have && (!single || (mask & 1u));
    const uint32_t anchor = mask & (~mask + 1u);
    const uint32_t rest = mask ^ anchor;
    uint32_t rl = 0, rr = rest;
#pragma unroll
    for (int q = 0; q < kLoBits; ++q) {
        const uint32_t b = rr & (~rr + 1u);
        rl |= b;
        rr ^= b;
    }
    const uint32_t rh = rest ^ rl;
    uint32_t slo = 0;
    {
        uint32_t bl = rl;
        int v = j;
#pragma unroll
        for (int q = 0; q < kLoBits; ++q) {
            const uint32_t b = bl & (~bl + 1u);
            if (v & 1) slo |= b;
            v >>= 1;
            bl ^= b;
        }
    }
    const uint32_t base_lo = anchor | slo;
    uint32_t shi = rh;
    double acc = rec ? a[mask] : 0.0;
    double rs = fabs(acc);
    const int steps = R / G;
    for (int st0 = 0; st0 < steps; st0 += kFoldSteps) {
        // every index below is a subset of the mask, so the loads are in
        // bounds for all lanes; invalid terms are replaced by 0
#pragma unroll
        for (int u = 0; u < kFoldSteps; ++u) {
            const int st = st0 + u;
            const uint32_t sub = base_lo | shi;
            const double prod = dp[sub] * a[mask ^ sub];
            const bool valid = rec && st < steps && !(st == 0 && j == G - 1);
            wtb[u * 64 + lane] = valid ? prod : 0.0;
            shi = (shi - 1u) & rh;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (rec && j == 0) {
            // the group's lane 0 folds the G terms of each step in descending
            // lane order, reading them as pairs (q - 1, q) sixteen bytes at a
            // time, the next eight terms in flight while eight fold
            constexpr int kChunks = G / 8;
#pragma unroll
            for (int u = 0; u < kFoldSteps; ++u) {
                const double2* src = reinterpret_cast<const double2*>(wtb + u * 64 + gi * G);
                double2 cur[4], nxt[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) cur[i] = src[G / 2 - 1 - i];
#pragma unroll
                for (int c = 0; c < kChunks; ++c) {
                    if (c + 1 < kChunks) {
#pragma unroll
                        for (int i = 0; i < 4; ++i) nxt[i] = src[G / 2 - 1 - 4 * (c + 1) - i];
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        acc = acc - cur[i].y;
                        rs += fabs(cur[i].y);
                        acc = acc - cur[i].x;
                        rs += fabs(cur[i].x);
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i) cur[i] = nxt[i];
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (have && j == 0) {
        if (rec) {
            if (filter && fabs(acc) < thr * rs) acc = 0.0;
            dp[mask] = acc;
            if (rhs) rhs[mask] = rs;
        } else {
            dp[mask] = a[mask];
        }
    }
}

// One popcount layer in tree order (the FP64 matrix-core and split-FP32
// modes, whose values are not bit-identical to the reference anyway): G = 2^lg
// lanes per mask, lane j of a group summing the terms x = G X + j of the
// chain enumeration of group_chain_round (s_lo the j-th subset of the lowest
// lg bits of the rest, s_hi descending over the subsets of the others), then
// a butterfly over the group's lanes.  acc = a[mask] - (that sum), rhs =
// |a[mask]| + (the sum of |terms|, alike), and the round-off filter of the
// reference (include/cdet/integrand_replica_randomly_sampled.h:330-341) on
// them.  The butterfly leaves the same bits in every lane of a group (a + b =
// b + a), so a value depends on (n, k, lg) only, never on the workgroup shape.
// Items q < classes M: class q / M at offset cls_stride, mask masks_by_pop[beg
// + q mod M]; the loop is uniform per wave (a group is all valid or all
// discarded), steps = R / G is a multiple of 8.  Per step a lane issues two
// LDS reads and seven VALU instructions; the serial chain of the reference
// order costs about as much per term on a single lane.
__device__ __forceinline__ void tree_layer(const double* a, double* dp, double* rhs, uint32_t cls_stride,
                                           int classes, int beg, int M, int R, int lg, double thr, bool filter,
                                           const uint16_t* masks_by_pop, bool single) {
    const int G = 1 << lg;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int waves = blockDim.x >> 6;
    const int per_wave = 64 >> lg;
    const int gi = lane >> lg;
    const int j = lane & (G - 1);
    const int items = classes * M;
    const int steps = R >> lg;
    for (int q0 = wave * per_wave; q0 < items; q0 += waves * per_wave) {
        const int q = q0 + gi;
        const bool have = q < items;
        const int cls = (have && q >= M) ? 1 : 0;
        const uint32_t off = cls ? cls_stride : 0u;
        // a discarded group reads the layer's first mask: its loads stay in bounds
        const uint32_t mask = masks_by_pop[beg + (have ? q - cls * M : 0)];
        const bool rec = have && (!single || (mask & 1u));
        const uint32_t anchor = mask & (~mask + 1u);
        const uint32_t rest = mask ^ anchor;
        uint32_t rl = 0, rr = rest;
        for (int b = 0; b < lg; ++b) {
            const uint32_t low = rr & (~rr + 1u);
            rl |= low;
            rr ^= low;
        }
        const uint32_t rh = rr;
        uint32_t slo = 0;
        {
            uint32_t bl = rl;
            int v = j;
            for (int b = 0; b < lg; ++b) {
                const uint32_t low = bl & (~bl + 1u);
                if (v & 1) slo |= low;
                v >>= 1;
                bl ^= low;
            }
        }
        const double* ac = a + off;
        double* dc = dp + off;
        // dp[anchor | s_lo | s_hi] and a[mask ^ sub] = a[(rest ^ s_lo) - s_hi]:
        // one add from a per-lane base, s_hi kept as a byte offset
        const char* dpb = reinterpret_cast<const char*>(dc + (anchor | slo));
        const char* ab = reinterpret_cast<const char*>(ac + (rest ^ slo));
        const uint32_t rh8 = rh << 3;
        uint32_t s8 = rh8;
        double part = 0.0, prs = 0.0;
        for (int st = 0; st < steps; st += 8) {
            double dv[8], av[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                dv[u] = *reinterpret_cast<const double*>(dpb + s8);
                av[u] = *reinterpret_cast<const double*>(ab - s8);
                s8 = (s8 - 8u) & rh8;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                double prod = dv[u] * av[u];
                // the first term of the last lane is sub = mask itself, no term
                if (u == 0 && st == 0 && j == G - 1) prod = 0.0;
                part += prod;
                prs += fabs(prod);
            }
        }
        for (int o = 1; o < G; o <<= 1) {
            part += __shfl_xor(part, o);
            prs += __shfl_xor(prs, o);
        }
        if (have && j == 0) {
            if (rec) {
                const double am = ac[mask];
                double acc = am - part;
                const double rs = fabs(am) + prs;
                if (filter && fabs(acc) < thr * rs) acc = 0.0;
                dc[mask] = acc;
                if (rhs) rhs[off + mask] = rs;
            } else {
                dc[mask] = ac[mask];
            }
        }
    }
}

// Lanes per mask (log2) of the tree-order form of popcount layer k at order n,
// for the workgroups that usually run it: both classes of orders up to 10 on
// 256 lanes, one class of orders 11, 12 on 256 and of 13, 14 on 1024.  Each
// choice G = 2^lg (at least 8 steps per lane) is priced as rounds x (steps x 9
// + lg x 8 + 30) instructions per lane, lg = 0 (the serial chain) as rounds x
// (9 (R - 1) + 20); the cheapest wins.  A fixed function of (n, k), so the
// values do not change with the launch shapes the tuning picks.
__host__ inline int tree_lanes_log2(int n, int k) {
    if (k < 2 || k > n) return 0;
    double masks = 1.0;
    for (int i = 0; i < k; ++i) masks = masks * (n - i) / (i + 1);
    const long long M = static_cast<long long>(masks + 0.5) * (n <= 10 ? 2 : 1);
    const long long L = n >= 13 ? 1024 : 256;
    const long long R = 1LL << (k - 1);
    long long best = ((M + L - 1) / L) * (9 * (R - 1) + 20);
    int best_lg = 0;
    for (int lg = 1; lg <= 6 && (R >> lg) >= 8; ++lg) {
        const long long G = 1LL << lg;
        const long long cost = ((M * G + L - 1) / L) * ((R >> lg) * 9 + lg * 8 + 30);
        if (cost < best) {
            best = cost;
            best_lg = lg;
        }
    }
    return best_lg;
}

template <int G>
__device__ __forceinline__ void group_chains(const double* a, double* dp, int beg, int end, int R, double thr,
                                             bool filter, const uint16_t* masks_by_pop, double* rhs, bool single,
                                             double* tb, int wave0 = 0, int nwaves = kFoldWaves) {
    static_assert(G == 8 || G == 16 || G == 32 || G == 64, "groups fold eight terms at a time");
    constexpr int kPerWave = 64 / G;
    const int tid = threadIdx.x;
    const int w = tid >> 6;
    const int wl = w - wave0;
    if (w >= kFoldWaves || wl < 0 || wl >= nwaves) return;  // uniform per wave
    // rounds of nwaves x kPerWave masks
    for (int base = beg + wl * kPerWave; base < end; base += nwaves * kPerWave)
        group_chain_round<G>(a, dp, base, end, R, thr, filter, masks_by_pop, rhs, single, tb + w * (kFoldSteps * 64));
}

// Subset recursion over all popcount layers, workgroup-cooperative.
//   all-subset (single = false; include/cdet/integrand_replica_randomly_sampled.h:292-347):
//     every mask recurses with anchor = its lowest bit;
//   single-anchor (single = true; :226-290): only masks containing bit 0
//     recurse (anchor = bit 0), the others keep dp = a.
// On entry a[mask] = pm_up * pm_down (a[0] = 0); on exit dp[mask] holds the
// connected value and, with rhs non-null, rhs[mask] the sum of |a[mask]| and
// |terms| of its chain (the round-off filter's scale, :330-341).  masks_by_pop
// lists every mask ordered by popcount, pop_off[k] the first index of
// popcount k; masks of one popcount are independent.  A layer runs one chain
// per lane (connected_chain) or as grouped chains (group_chains, see
// layer_group).  tb holds kFoldWaves x kFoldSteps x 64 doubles.
__device__ __forceinline__ void ztransform_layers(const double* a, double* dp, int n, double thr, bool filter,
                                         const uint16_t* masks_by_pop, const int* pop_off, double* rhs, bool single,
                                         double* tb, int g16 = 64, int g64 = 128, uint64_t tree_bits = 0) {
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const uint32_t sc = 1u << n;
    for (uint32_t m = tid; m < sc; m += nt) dp[m] = (m != 0 && (m & (m - 1)) == 0) ? a[m] : 0.0;
    __syncthreads();
    for (int bits = 2; bits <= n; ++bits) {
        const int beg = pop_off[bits];
        const int end = pop_off[bits + 1];
        const int R = 1 << (bits - 1);
        const int lg = static_cast<int>((tree_bits >> (3 * bits)) & 7u);
        const int G = lg ? -1 : layer_group(end - beg, R, nt >> 6, g16, g64);
        if (lg) {
            tree_layer(a, dp, rhs, 0u, 1, beg, end - beg, R, lg, thr, filter, masks_by_pop, single);
        } else if (G == 64) {
            group_chains<64>(a, dp, beg, end, R, thr, filter, masks_by_pop, rhs, single, tb);
        } else if (G == 16) {
            group_chains<16>(a, dp, beg, end, R, thr, filter, masks_by_pop, rhs, single, tb);
        } else {
            for (int it = beg + tid; it < end; it += nt) {
                const uint32_t mask = masks_by_pop[it];
                if (single && !(mask & 1u)) {
                    dp[mask] = a[mask];
                    continue;
                }
                const uint32_t anchor = mask & (~mask + 1u);
                double r;
                dp[mask] = connected_chain(a, dp, mask, anchor, mask ^ anchor, R - 1, thr, filter, r);
                if (rhs) rhs[mask] = r;
            }
        }
        __syncthreads();
    }
}

// All-subset recursion of both table classes of one configuration at once:
// a, dp (and rhs, if non-null) hold class 0 at [0, 2^n) and class 1 at
// [2^n, 2^(n+1)).  Each popcount layer runs the masks of both classes before
// one barrier: one lane per (class, mask), or grouped chains with the two
// classes' masks on different waves -- a single long chain per class on a
// wave each, up to 16 masks per class on 8-lane groups with two waves per
// class -- or the classes one after the other with all fold waves.  Every
// value is formed by the same chain as in ztransform_layers, so the results
// are identical to two separate runs.
__device__ __forceinline__ void ztransform_layers_pair(const double* a, double* dp, int n, double thr, bool filter,
                                                       const uint16_t* masks_by_pop, const int* pop_off, double* rhs,
                                                       double* tb, int g16 = 64, int g64 = 128,
                                                       uint64_t tree_bits = 0) {
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const int waves = nt >> 6;
    const uint32_t sc = 1u << n;
    for (uint32_t m = tid; m < 2 * sc; m += nt) {
        const uint32_t mm = m & (sc - 1);
        dp[m] = (mm != 0 && (mm & (mm - 1)) == 0) ? a[m] : 0.0;
    }
    __syncthreads();
    for (int bits = 2; bits <= n; ++bits) {
        const int beg = pop_off[bits];
        const int end = pop_off[bits + 1];
        const int M = end - beg;
        const int R = 1 << (bits - 1);
        const int lg = static_cast<int>((tree_bits >> (3 * bits)) & 7u);
        const int G = lg ? -1 : layer_group(M, R, waves, g16, g64);
        if (lg) {
            tree_layer(a, dp, rhs, sc, 2, beg, M, R, lg, thr, filter, masks_by_pop, false);
        } else if (G == 64 && waves >= 2) {
            // one long chain per class: class c on wave c
            group_chains<64>(a, dp, beg, end, R, thr, filter, masks_by_pop, rhs, false, tb, 0, 1);
            group_chains<64>(a + sc, dp + sc, beg, end, R, thr, filter, masks_by_pop, rhs ? rhs + sc : nullptr, false,
                             tb, 1, 1);
        } else if (G == 64) {
            group_chains<64>(a, dp, beg, end, R, thr, filter, masks_by_pop, rhs, false, tb);
            group_chains<64>(a + sc, dp + sc, beg, end, R, thr, filter, masks_by_pop, rhs ? rhs + sc : nullptr, false,
                             tb);
        } else if (G == 16 && waves >= 4 && M <= 16) {
            // both classes in one round: 8 lanes per mask, class c on waves
            // 2c and 2c + 1 (16 groups per class)
            group_chains<8>(a, dp, beg, end, R, thr, filter, masks_by_pop, rhs, false, tb, 0, 2);
            group_chains<8>(a + sc, dp + sc, beg, end, R, thr, filter, masks_by_pop, rhs ? rhs + sc : nullptr, false,
                            tb, 2, 2);
        } else if (G == 16) {
            group_chains<16>(a, dp, beg, end, R, thr, filter, masks_by_pop, rhs, false, tb);
            group_chains<16>(a + sc, dp + sc, beg, end, R, thr, filter, masks_by_pop, rhs ? rhs + sc : nullptr, false,
                             tb);
        } else {
            for (int q = tid; q < 2 * M; q += nt) {
                const int cls = q >= M ? 1 : 0;
                const uint32_t mask = masks_by_pop[beg + q - cls * M];
                const uint32_t anchor = mask & (~mask + 1u);
                const double* ac = a + cls * sc;
                double* dc = dp + cls * sc;
                double r;
                dc[mask] = connected_chain(ac, dc, mask, anchor, mask ^ anchor, R - 1, thr, filter, r);
                if (rhs) rhs[cls * sc + mask] = r;
            }
        }
        __syncthreads();
    }
}

}  // namespace cdet

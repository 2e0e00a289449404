// cdet_mfma.hpp -- matrix-core (MFMA) coefficient x basis contraction for the
// batched table evaluation.
//
// For the Q queries (u_q, v_q) that hit one coefficient block C (nu x nv):
//     R[a][q] = sum_b C[a][b] T_b(v_q)            (MFMA, M = a, N = q, K = b)
//     G_q     = sum_a T_a(u_q) R[a][q]            (per-lane FMA + cross-lane sum)
// One wave owns a tile of up to 32 queries (two 16-wide N tiles) and sweeps
// the block in 16-row slabs, so only two 16x16 accumulators are live; the
// A fragment of each K step is loaded once and used for both N tiles.
//
// Operand layouts (lane l, g = l / 16, j = l % 16):  A[j][g] and B[g][j] for
// both instructions; the accumulator register r of lane l holds
//   v_mfma_f32_16x16x4_f32:  D[4 g + r][j]   (four consecutive rows)
//   v_mfma_f64_16x16x4_f64:  D[4 r + g][j]   (rows strided by four)
//
// Chebyshev bases are generated in registers:  T_{k+4} = 2 T_4 T_k - T_{k-4}
// along K (each lane owns k = 4s + g) and T_{a+16} = 2 T_16 T_a - T_{a-16}
// across row slabs, seeded from T_0..T_16 by the ordinary three-term
// recurrence (T_{-m} = T_m).
//
// Modes:
//   MFMA64   FP64 operands, FP64 accumulation (v_mfma_f64_16x16x4_f64).
//   SPLIT32  C = C_hi + C_lo, B = B_hi + B_lo in FP32; the three products
//            hi*hi + hi*lo + lo*hi are accumulated in FP32 over one K step
//            (4 terms) and added into FP64 accumulators.
#pragma once

#include "cdet_scalars.hpp"

namespace cdet {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaTile = 32;  // queries per wave

// T_m(x) for m = 0..16.
__device__ inline void cheb_0_16(double x, double* t) {
    t[0] = 1.0;
    t[1] = x;
    const double tx = 2.0 * x;
#pragma unroll
    for (int m = 2; m <= 16; ++m) t[m] = tx * t[m - 1] - t[m - 2];
}

// Select a_g (g in 0..3) with two bit tests, branch-free.
__device__ inline double sel4(int g, double a0, double a1, double a2, double a3) {
    const double lo = (g & 1) ? a1 : a0;
    const double hi = (g & 1) ? a3 : a2;
    return (g & 2) ? hi : lo;
}

struct BasisK {  // lane's T_k(v) along K with k = 4s + g
    double cur, prev, w;
    __device__ inline void init(double v, int g) {
        const double t2 = 2.0 * v * v - 1.0;
        const double t3 = 2.0 * v * t2 - v;
        const double t4 = 2.0 * v * t3 - t2;
        cur = sel4(g, 1.0, v, t2, t3);
        prev = sel4(g, t4, t3, t2, v);  // T_{4-g} = T_{g-4}
        w = 2.0 * t4;
    }
    __device__ inline void step() {
        const double nx = __builtin_fma(w, cur, -prev);
        prev = cur;
        cur = nx;
    }
};

// Row of accumulator register r in slab m: 16 m + row_in_slab(g, r).
template <int MODE>
__device__ inline int row_in_slab(int g, int r) {
    return MODE == CDET_MODE_MFMA64 ? 4 * r + g : 4 * g + r;
}

template <int MODE>
struct BasisA {  // lane's T_a(u) for its four rows a = 16 m + row_in_slab(g, r)
    double cur[4], prev[4], w;
    __device__ inline void init(double u, int g) {
        double t[17];
        cheb_0_16(u, t);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (MODE == CDET_MODE_MFMA64) {  // rows g + 4 r
                cur[r] = sel4(g, t[4 * r], t[4 * r + 1], t[4 * r + 2], t[4 * r + 3]);
                prev[r] = sel4(g, t[16 - 4 * r], t[15 - 4 * r], t[14 - 4 * r], t[13 - 4 * r]);
            } else {  // rows 4 g + r
                cur[r] = sel4(g, t[r], t[4 + r], t[8 + r], t[12 + r]);
                prev[r] = sel4(g, t[16 - r], t[12 - r], t[8 - r], t[4 - r]);
            }
        }
        w = 2.0 * t[16];
    }
    __device__ inline void step() {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const double nx = __builtin_fma(w, cur[r], -prev[r]);
            prev[r] = cur[r];
            cur[r] = nx;
        }
    }
};

// Work split for the persistent contraction kernels.  Workgroups are dealt
// round-robin to the 8 XCDs (workgroup b runs on XCD b % 8), so the tiles are
// cut into 8 equal contiguous ranges, one per XCD: every XCD gets the same
// share and the tiles of one coefficient block (consecutive) share an L2.
// Within its XCD a wave strides over the range.
struct TileRange {
    int first, end, stride;
};
__device__ inline TileRange xcd_tiles(int total, int wave, int wpb) {
    const int grid = gridDim.x;
    if (grid % 8 != 0) return TileRange{static_cast<int>(blockIdx.x) * wpb + wave, total, grid * wpb};
    const int x = blockIdx.x & 7;
    const int chunk = (total + 7) >> 3;
    const int end = min(total, (x + 1) * chunk);
    return TileRange{x * chunk + static_cast<int>(blockIdx.x >> 3) * wpb + wave, end, (grid >> 3) * wpb};
}

template <int MODE>
__global__ __launch_bounds__(256) void k_contract_mfma(KScalars ks, const int32_t* __restrict__ n_tiles,
                                                       const int32_t* __restrict__ tile_key,
                                                       const int32_t* __restrict__ tile_start,
                                                       const int32_t* __restrict__ tile_len,
                                                       const double* __restrict__ s_u, const double* __restrict__ s_v,
                                                       const int32_t* __restrict__ s_dest, double* __restrict__ g_up,
                                                       double* __restrict__ g_dn) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4;
    const int j = lane & 15;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const TileRange tr = xcd_tiles(*n_tiles, wave, blockDim.x >> 6);
    const int nu = ks.nu, nv = ks.nv;
    const int mslabs = (nu + 15) >> 4;
    const int ksteps = (nv + 3) >> 2;
    for (int t = tr.first; t < tr.end; t += tr.stride) {
        const int tile = __builtin_amdgcn_readfirstlane(t);
        const int32_t key = __builtin_amdgcn_readfirstlane(tile_key[tile]);
        const int32_t start = __builtin_amdgcn_readfirstlane(tile_start[tile]);
        const int32_t len = __builtin_amdgcn_readfirstlane(tile_len[tile]);
        const double* cu;
        const double* cd;
        bool same;
        ks.key_tables(key, cu, cd, same);
        const float* hu = nullptr;
        const float* lu = nullptr;
        const float* hd = nullptr;
        const float* ld = nullptr;
        if (MODE == CDET_MODE_SPLIT32) ks.key_split(key, hu, lu, hd, ld);
        // queries of the two N tiles owned by this lane
        const int q0 = j, q1 = j + 16;
        const int p0 = start + (q0 < len ? q0 : 0);
        const int p1 = start + (q1 < len ? q1 : 0);
        const double u0 = s_u[p0], v0 = s_v[p0];
        const double u1 = s_u[p1], v1 = s_v[p1];
        BasisA<MODE> A0, A1;
        A0.init(u0, g);
        A1.init(u1, g);
        double G0u = 0.0, G1u = 0.0, G0d = 0.0, G1d = 0.0;
        for (int ms = 0; ms < mslabs; ++ms) {
            const int row = 16 * ms + j;
            const bool row_ok = row < nu;
            d4 acc0u = {0.0, 0.0, 0.0, 0.0}, acc1u = acc0u, acc0d = acc0u, acc1d = acc0u;
            BasisK B0, B1;
            B0.init(v0, g);
            B1.init(v1, g);
            const int64_t roff = static_cast<int64_t>(row_ok ? row : 0) * nv;
            for (int s = 0; s < ksteps; ++s) {
                const int k = 4 * s + g;
                const bool ok = row_ok && k < nv;
                if (MODE == CDET_MODE_MFMA64) {
                    const double au = ok ? cu[roff + k] : 0.0;
                    acc0u = __builtin_amdgcn_mfma_f64_16x16x4f64(au, B0.cur, acc0u, 0, 0, 0);
                    acc1u = __builtin_amdgcn_mfma_f64_16x16x4f64(au, B1.cur, acc1u, 0, 0, 0);
                    if (!same) {
                        const double ad = ok ? cd[roff + k] : 0.0;
                        acc0d = __builtin_amdgcn_mfma_f64_16x16x4f64(ad, B0.cur, acc0d, 0, 0, 0);
                        acc1d = __builtin_amdgcn_mfma_f64_16x16x4f64(ad, B1.cur, acc1d, 0, 0, 0);
                    }
                } else {
                    const float ah = ok ? hu[roff + k] : 0.0f;
                    const float al = ok ? lu[roff + k] : 0.0f;
                    const float b0h = static_cast<float>(B0.cur);
                    const float b0l = static_cast<float>(B0.cur - static_cast<double>(b0h));
                    const float b1h = static_cast<float>(B1.cur);
                    const float b1l = static_cast<float>(B1.cur - static_cast<double>(b1h));
                    const f4 z = {0.0f, 0.0f, 0.0f, 0.0f};
                    f4 t0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, b0h, z, 0, 0, 0);
                    t0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, b0l, t0, 0, 0, 0);
                    t0 = __builtin_amdgcn_mfma_f32_16x16x4f32(al, b0h, t0, 0, 0, 0);
                    f4 t1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, b1h, z, 0, 0, 0);
                    t1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, b1l, t1, 0, 0, 0);
                    t1 = __builtin_amdgcn_mfma_f32_16x16x4f32(al, b1h, t1, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        acc0u[r] += static_cast<double>(t0[r]);
                        acc1u[r] += static_cast<double>(t1[r]);
                    }
                    if (!same) {
                        const float dh = ok ? hd[roff + k] : 0.0f;
                        const float dl = ok ? ld[roff + k] : 0.0f;
                        f4 e0 = __builtin_amdgcn_mfma_f32_16x16x4f32(dh, b0h, z, 0, 0, 0);
                        e0 = __builtin_amdgcn_mfma_f32_16x16x4f32(dh, b0l, e0, 0, 0, 0);
                        e0 = __builtin_amdgcn_mfma_f32_16x16x4f32(dl, b0h, e0, 0, 0, 0);
                        f4 e1 = __builtin_amdgcn_mfma_f32_16x16x4f32(dh, b1h, z, 0, 0, 0);
                        e1 = __builtin_amdgcn_mfma_f32_16x16x4f32(dh, b1l, e1, 0, 0, 0);
                        e1 = __builtin_amdgcn_mfma_f32_16x16x4f32(dl, b1h, e1, 0, 0, 0);
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            acc0d[r] += static_cast<double>(e0[r]);
                            acc1d[r] += static_cast<double>(e1[r]);
                        }
                    }
                }
                B0.step();
                B1.step();
            }
            // Row slab contribution: rows 16 ms + row_in_slab(g, r) of this lane.
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const bool ok = (16 * ms + row_in_slab<MODE>(g, r)) < nu;
                const double w0 = ok ? A0.cur[r] : 0.0;
                const double w1 = ok ? A1.cur[r] : 0.0;
                G0u = __builtin_fma(w0, acc0u[r], G0u);
                G1u = __builtin_fma(w1, acc1u[r], G1u);
                if (!same) {
                    G0d = __builtin_fma(w0, acc0d[r], G0d);
                    G1d = __builtin_fma(w1, acc1d[r], G1d);
                }
            }
            A0.step();
            A1.step();
        }
        // Sum the four row groups (lanes j, j+16, j+32, j+48).
        G0u += __shfl_xor(G0u, 16);
        G0u += __shfl_xor(G0u, 32);
        G1u += __shfl_xor(G1u, 16);
        G1u += __shfl_xor(G1u, 32);
        if (!same) {
            G0d += __shfl_xor(G0d, 16);
            G0d += __shfl_xor(G0d, 32);
            G1d += __shfl_xor(G1d, 16);
            G1d += __shfl_xor(G1d, 32);
        }
        if (g == 0) {
            if (q0 < len) {
                const int32_t d0 = s_dest[start + q0];
                g_up[d0] = G0u;
                g_dn[d0] = same ? G0u : G0d;
            }
            if (q1 < len) {
                const int32_t d1 = s_dest[start + q1];
                g_up[d1] = G1u;
                g_dn[d1] = same ? G1u : G1d;
            }
        }
    }
}

typedef double d4v __attribute__((ext_vector_type(4)));

constexpr int kF64Tile = 64;  // queries per wave in k_contract_f64

// Coefficient steps in flight per wave in contract_f64_tile (a third set
// measured no faster).
template <int NT>
constexpr int contract_f64_prefetch() {
    return 2;
}

// One tile of k_contract_f64 with NT live 16-wide N tiles (16 (NT-1) < len <=
// 16 NT).  K is visited in groups of 16; MFMA step r of group t uses
// k = 16 t + 4 r + g on both operands, so each lane's K basis is one stride-4
// Chebyshev sequence (BasisK).  The coefficients are read from the
// step-order copy of the table (k_step_order): the four A values of a step
// are four contiguous 512-byte wave reads.  Accumulator register r holds row
// 16 m + 4 r + g, again a stride-4 sequence (BasisK over u).  The next group's
// coefficients are loaded before the current group's MFMAs; every N tile
// reuses each A fragment.
template <bool SAME, int NT>
__device__ inline void contract_f64_tile(const KScalars& ks, int g, int j, const double* __restrict__ cu,
                                         const double* __restrict__ cd, int32_t start, int32_t len,
                                         const double* __restrict__ s_u, const double* __restrict__ s_v,
                                         const int32_t* __restrict__ s_dest, double* __restrict__ g_up,
                                         double* __restrict__ g_dn) {
    const int nu = ks.nu, nv = ks.nv;
    const int mslabs = nu >> 4;
    const int groups = nv >> 4;
    BasisK Au[NT];
    double v[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int q = 16 * nt + j;
        const int p = start + (q < len ? q : 0);
        v[nt] = s_v[p];
        Au[nt].init(s_u[p], g);
    }
    double G[NT], Gd[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) G[nt] = Gd[nt] = 0.0;
    // The (slab, group) pairs form one stream of mslabs * groups steps, 256
    // values each in step order; the coefficients of later steps are loaded
    // (rotating registers, fenced from the scheduler) while step k runs.
    const int total = mslabs * groups;
    int64_t loff = 16 * g + j;  // next load: step lk, value 64 r + lane
    int lk = 0;
    auto load_next = [&](double* x, double* y) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            x[r] = cu[loff + 64 * r];
            if (!SAME) y[r] = cd[loff + 64 * r];
        }
        if (lk + 1 < total) {  // past the end the last step is read again
            ++lk;
            loff += 256;
        }
    };
    d4 acc[NT], accd[NT];
    BasisK Bv[NT];
    int ct = 0;
    auto compute = [&](const double* x, const double* y) {
        if (ct == 0) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                acc[nt] = d4{0.0, 0.0, 0.0, 0.0};
                accd[nt] = acc[nt];
                Bv[nt].init(v[nt], g);
            }
        }
        // Four recurrence steps per lane and N tile with the two basis
        // registers alternating roles (no register moves): steps r = 0, 2 read
        // cur and overwrite prev with T_{k+4}; r = 1, 3 read prev and
        // overwrite cur, leaving (cur, prev) = (T_{k+16}, T_{k+12}).
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                double& rd = (r & 1) ? Bv[nt].prev : Bv[nt].cur;
                double& wr = (r & 1) ? Bv[nt].cur : Bv[nt].prev;
                acc[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[r], rd, acc[nt], 0, 0, 0);
                if (!SAME) accd[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(y[r], rd, accd[nt], 0, 0, 0);
                wr = __builtin_fma(Bv[nt].w, rd, -wr);
            }
        }
        if (++ct == groups) {  // slab done: rows 16 m + 4 r + g of the accumulators
            ct = 0;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    G[nt] = __builtin_fma(Au[nt].cur, acc[nt][r], G[nt]);
                    if (!SAME) Gd[nt] = __builtin_fma(Au[nt].cur, accd[nt][r], Gd[nt]);
                    Au[nt].step();
                }
            }
        }
    };
    // PF register sets: step k + PF - 1 is loaded while step k runs.
    constexpr int PF = contract_f64_prefetch<NT>();
    double a[PF][4], b[PF][4];
#pragma unroll
    for (int d = 0; d + 1 < PF; ++d) load_next(a[d], b[d]);
    for (int k = 0; k < total; k += PF) {
#pragma unroll
        for (int d = 0; d < PF; ++d) {
            load_next(a[(d + PF - 1) % PF], b[(d + PF - 1) % PF]);
            __builtin_amdgcn_sched_barrier(0);
            if (k + d < total) compute(a[d], b[d]);
        }
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        double x = G[nt];
        x += __shfl_xor(x, 16);
        x += __shfl_xor(x, 32);
        double y = x;
        if (!SAME) {
            y = Gd[nt];
            y += __shfl_xor(y, 16);
            y += __shfl_xor(y, 32);
        }
        const int q = 16 * nt + j;
        if (g == 0 && q < len) {
            const int32_t d = s_dest[start + q];
            g_up[d] = x;
            g_dn[d] = y;
        }
    }
}

// One tile of at most 16 queries with the K groups outside the row slabs: the
// slabs are taken S at a time (S live accumulators) and, within such a chunk,
// the basis values T_k(v) of a K group (k = 16 t + 4 r + g) are formed once
// and feed the products of all S slabs.  FP64 vector instructions hold up the
// FP64 matrix core of their SIMD (one FP64 FMA per 16x16x4 product costs
// about 15% of the product rate), so the basis work falls from one recurrence
// step per product (contract_f64_tile<SAME, 1>) to one per S products.  The
// products of a slab and their order of accumulation are those of
// contract_f64_tile<SAME, 1> (slab-major over the step-order table, K groups
// in order), and so are the basis values and the row fold: the values are the
// same bit for bit.  Coefficient offsets are wave-uniform, so the loads take
// a scalar base and a fixed per-lane offset; the step PF - 1 ahead of the one
// computed is loaded into a rotating register set.
template <bool SAME, int S>
__device__ inline void contract_f64_tile_slabs(const KScalars& ks, int g, int j, const double* __restrict__ cu,
                                               const double* __restrict__ cd, int32_t start, int32_t len,
                                               const double* __restrict__ s_u, const double* __restrict__ s_v,
                                               const int32_t* __restrict__ s_dest, double* __restrict__ g_up,
                                               double* __restrict__ g_dn) {
    constexpr int PF = 2;
    static_assert(S % PF == 0, "the rotating register sets must divide the chunk");
    const int groups = ks.nv >> 4;
    const int nchunk = (ks.nu >> 4) / S;
    const int p = start + (j < len ? j : 0);
    const double v = s_v[p];
    BasisK Au;
    Au.init(s_u[p], g);
    double G = 0.0, Gd = 0.0;
    // byte offset of the lane's element within a step: the loads take a
    // wave-uniform 64-bit base, this unsigned 32-bit offset and an immediate
    // 512 r, with no per-step vector address arithmetic
    const uint32_t lane_byte = static_cast<uint32_t>(16 * g + j) * 8u;
    double a[PF][4], b[PF][4];
    auto load = [&](double* x, double* y, int off) {
        const char* bu = reinterpret_cast<const char*>(cu + __builtin_amdgcn_readfirstlane(off));
        const char* bd = reinterpret_cast<const char*>(cd + __builtin_amdgcn_readfirstlane(off));
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            x[r] = *reinterpret_cast<const double*>(bu + lane_byte + 512 * r);
            if (!SAME) y[r] = *reinterpret_cast<const double*>(bd + lane_byte + 512 * r);
        }
    };
    const int slab_stride = groups * 256;  // doubles between the steps of one K group in two slabs
#pragma unroll
    for (int d = 0; d + 1 < PF; ++d) load(a[d], b[d], d * slab_stride);  // steps (chunk 0, group 0, slab d)
    for (int c = 0; c < nchunk; ++c) {
        d4 acc[S], accd[S];
#pragma unroll
        for (int mm = 0; mm < S; ++mm) acc[mm] = accd[mm] = d4{0.0, 0.0, 0.0, 0.0};
        BasisK Bv;
        Bv.init(v, g);
        for (int t = 0; t < groups; ++t) {
            // T_k(v) for k = 16 t + 4 r + g, by the recurrence of contract_f64_tile
            double bv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                double& rd = (r & 1) ? Bv.prev : Bv.cur;
                double& wr = (r & 1) ? Bv.cur : Bv.prev;
                bv[r] = rd;
                wr = __builtin_fma(Bv.w, rd, -wr);
            }
#pragma unroll
            for (int mm = 0; mm < S; ++mm) {
                // the step PF - 1 after (c, t, mm); past the last step the last
                // step is read again
                int lm = mm + PF - 1, lt = t, lc = c;
                if (lm >= S) {
                    lm -= S;
                    if (++lt == groups) {
                        lt = 0;
                        ++lc;
                    }
                }
                if (lc == nchunk) {
                    lc = nchunk - 1;
                    lt = groups - 1;
                    lm = S - 1;
                }
                load(a[(mm + PF - 1) % PF], b[(mm + PF - 1) % PF], (lc * S + lm) * slab_stride + lt * 256);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    acc[mm] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mm % PF][r], bv[r], acc[mm], 0, 0, 0);
                    if (!SAME)
                        accd[mm] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[mm % PF][r], bv[r], accd[mm], 0, 0, 0);
                }
            }
        }
        // rows 16 (S c + mm) + 4 r + g of the accumulators, slabs in order
#pragma unroll
        for (int mm = 0; mm < S; ++mm) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                G = __builtin_fma(Au.cur, acc[mm][r], G);
                if (!SAME) Gd = __builtin_fma(Au.cur, accd[mm][r], Gd);
                Au.step();
            }
        }
    }
    G += __shfl_xor(G, 16);
    G += __shfl_xor(G, 32);
    double y = G;
    if (!SAME) {
        y = Gd;
        y += __shfl_xor(y, 16);
        y += __shfl_xor(y, 32);
    }
    if (g == 0 && j < len) {
        const int32_t d = s_dest[start + j];
        g_up[d] = G;
        g_dn[d] = y;
    }
}

// Slabs per chunk of contract_f64_tile_slabs for nu = 16 mslabs: 2 where
// mslabs is even, else 0 (the slab-major form contract_f64_tile<SAME, 1>
// runs).  Order-10 batches of the 96x96 tables: 65.4 us per contraction with
// chunks of 2 slabs (66 registers, seven waves per SIMD), 67.4 us with chunks
// of 6 (106 registers, four waves), 70.0 us slab-major.
__host__ __device__ constexpr int contract_f64_slab_chunk(int mslabs, bool same) {
    return (void)same, mslabs % 2 == 0 ? 2 : 0;
}
template <bool SAME, int S>
constexpr int contract_f64_slab_waves() {
    return SAME ? 7 : 4;
}
template <bool SAME, int S>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(contract_f64_slab_waves<SAME, S>()))) void
k_contract_f64_slabs(KScalars ks, const int32_t* __restrict__ n_tiles, const int32_t* __restrict__ tile_key,
                     const int32_t* __restrict__ tile_start, const int32_t* __restrict__ tile_len,
                     const double* __restrict__ s_u, const double* __restrict__ s_v,
                     const int32_t* __restrict__ s_dest, double* __restrict__ g_up, double* __restrict__ g_dn) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4;
    const int j = lane & 15;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const TileRange tr = xcd_tiles(*n_tiles, wave, blockDim.x >> 6);
    for (int t0 = tr.first; t0 < tr.end; t0 += tr.stride) {
        const int tile = __builtin_amdgcn_readfirstlane(t0);
        const int32_t key = __builtin_amdgcn_readfirstlane(tile_key[tile]);
        const int32_t start = __builtin_amdgcn_readfirstlane(tile_start[tile]);
        const int32_t len = __builtin_amdgcn_readfirstlane(tile_len[tile]);
        const double* cu;
        const double* cd;
        bool same;
        ks.key_steps(key, cu, cd, same);
        contract_f64_tile_slabs<SAME, S>(ks, g, j, cu, cd, start, len, s_u, s_v, s_dest, g_up, g_dn);
    }
}

// FP64 matrix-core contraction for nu % 16 == 0 and nv % 16 == 0: one wave
// per tile of up to 64 queries of one coefficient block (tiles split over the
// XCDs by xcd_tiles).  SAME selects the form for contexts whose up and down
// tables are one table for both classes; it is held to 128 registers so that
// four waves share each SIMD.
// MAXNT = 1 is the form for tiles of at most 16 queries (one N tile): it
// needs a quarter of the accumulators, so that eight waves (six with separate
// up and down tables) share each SIMD without spilling.
template <bool SAME, int MAXNT>
constexpr int contract_f64_waves() {
    return MAXNT == 1 ? (SAME ? 8 : 6) : (SAME ? 4 : 2);
}
template <bool SAME, int MAXNT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(contract_f64_waves<SAME, MAXNT>()))) void
k_contract_f64(
    KScalars ks, const int32_t* __restrict__ n_tiles, const int32_t* __restrict__ tile_key,
    const int32_t* __restrict__ tile_start, const int32_t* __restrict__ tile_len, const double* __restrict__ s_u,
    const double* __restrict__ s_v, const int32_t* __restrict__ s_dest, double* __restrict__ g_up,
    double* __restrict__ g_dn) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4;
    const int j = lane & 15;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const TileRange tr = xcd_tiles(*n_tiles, wave, blockDim.x >> 6);
    for (int t0 = tr.first; t0 < tr.end; t0 += tr.stride) {
        const int tile = __builtin_amdgcn_readfirstlane(t0);
        const int32_t key = __builtin_amdgcn_readfirstlane(tile_key[tile]);
        const int32_t start = __builtin_amdgcn_readfirstlane(tile_start[tile]);
        const int32_t len = __builtin_amdgcn_readfirstlane(tile_len[tile]);
        const double* cu;
        const double* cd;
        bool same;
        ks.key_steps(key, cu, cd, same);
        const int nts = (len + 15) >> 4;
        if (MAXNT == 1)
            contract_f64_tile<SAME, 1>(ks, g, j, cu, cd, start, len, s_u, s_v, s_dest, g_up, g_dn);
        else if (nts == 4)
            contract_f64_tile<SAME, 4>(ks, g, j, cu, cd, start, len, s_u, s_v, s_dest, g_up, g_dn);
        else if (nts == 3)
            contract_f64_tile<SAME, 3>(ks, g, j, cu, cd, start, len, s_u, s_v, s_dest, g_up, g_dn);
        else if (nts == 2)
            contract_f64_tile<SAME, 2>(ks, g, j, cu, cd, start, len, s_u, s_v, s_dest, g_up, g_dn);
        else
            contract_f64_tile<SAME, 1>(ks, g, j, cu, cd, start, len, s_u, s_v, s_dest, g_up, g_dn);
    }
}

}  // namespace cdet

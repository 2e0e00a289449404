// cdet_kernels.hip -- gfx950 kernels of the CDet integrand path and the
// C-ABI declared in include/cdet_hip.h.
//
// Pipeline of one batched evaluation (all configurations packed, each owning
// 2^n output doubles):
//   k_plan           one workgroup per configuration: decode vertices, build
//                    the 2*n^2 lookup descriptors (twisted and ordinary
//                    class), resolve corner lookups directly, and enqueue the
//                    rest by coefficient block (class, branch, sitepair)
//   k_scan_keys      one workgroup: block offsets and per-block query tiles
//   k_scatter        queries grouped by coefficient block
//   k_contract_*     persistent: each wave takes a tile of queries sharing one
//                    coefficient block, so the block is read once per tile
//   k_minors         order-bucketed: one workgroup per (configuration, table
//                    class), vertex matrices, principal minors and subset
//                    recursion in LDS; k_minors_repair finishes the rare
//                    pivot fallbacks; k_combine_* forms twisted - 2*ordinary
// The plain form (one configuration per launch) uses k_pairs_plain (one wave
// per Green pair) and the same k_minors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <condition_variable>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cdet_chunks.h"
#include "../../include/cdet_hip.h"
#include "../../include/cdet_mcmcmc.h"
#include "cdet_bound.hpp"
#include "cdet_cdet.hpp"
#include "cdet_fpm_wave.hpp"
#include "cdet_device.hpp"
#include "cdet_scalars.hpp"
#include "cdet_mfma.hpp"

using namespace cdet;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess) return fail(CDET_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

struct DevTable {
    double* upper = nullptr;
    double* lower = nullptr;
    double* ucorner = nullptr;
    double* lcorner = nullptr;
    float* split[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [branch][hi, lo]
    double* steps[2] = {nullptr, nullptr};  // [branch] coefficients in matrix-core step order
    double* abssum[2] = {nullptr, nullptr};  // [branch] sum of |c_ab| per coefficient block
    double* rowsum = nullptr;  // lower branch: per block and row a, sum over b of c_ab (lookups at v = 1)
    int alias = -1;  // kind whose storage this kind shares
};

}  // namespace

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Plain form: one wave per Green pair of one configuration.  Lanes own rows a
// of the coefficient block (row Clenshaw over v); lane 0 then runs the
// Clenshaw over u in the reference order.
__global__ __launch_bounds__(64) void k_pairs_plain(KScalars ks, const double* __restrict__ verts, int n,
                                                    double* __restrict__ g_up, double* __restrict__ g_dn) {
    __shared__ double rows_up[256];
    __shared__ double rows_dn[256];
    const int pair = blockIdx.x;  // cls * n^2 + idx
    const int n2 = n * n;
    const int cls = pair / n2;
    const int idx = pair - cls * n2;
    const int i = idx / n;
    const int j = idx - i * n;
    const int rc = cls == 0 ? 2 : 1;
    const Vtx vi = decode_vertex(verts + 4 * i, ks.beta, ks.N1, ks.N2, rc);
    const Vtx vj = decode_vertex(verts + 4 * j, ks.beta, ks.N1, ks.N2, rc);
    const bool upper = (i != j) && (vi.tau_unit > vj.tau_unit);
    const Desc d = make_desc(vi, vj, upper, ks.beta, ks.branch_eps, ks.N1, ks.N2);
    const int ku = cls == 0 ? CDET_TABLE_TWISTED_UP : CDET_TABLE_ORDINARY_UP;
    const int kd = ku + 1;
    const int br = upper ? 0 : 1;
    const bool same = ks.same_spin[cls] != 0;
    if (d.corner) {
        if (threadIdx.x == 0) {
            g_up[pair] = ks.corner[ku][br][d.sitepair];
            g_dn[pair] = ks.corner[kd][br][d.sitepair];
        }
        return;
    }
    const int64_t base = d.sitepair * static_cast<int64_t>(ks.nu) * ks.nv;
    const double* cu = ks.coeff[ku][br] + base;
    const double* cd = ks.coeff[kd][br] + base;
    for (int a = threadIdx.x; a < ks.nu; a += 64) {
        rows_up[a] = clenshaw_row(cu + static_cast<int64_t>(a) * ks.nv, ks.nv - 1, d.v);
        if (!same) rows_dn[a] = clenshaw_row(cd + static_cast<int64_t>(a) * ks.nv, ks.nv - 1, d.v);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double gu = clenshaw_row(rows_up, ks.nu - 1, d.u);
        g_up[pair] = gu;
        g_dn[pair] = same ? gu : clenshaw_row(rows_dn, ks.nu - 1, d.u);
    }
}

// Batched planning: one workgroup per configuration.
//   cfg_order[c], cfg_vbase[c] (first vertex), cfg_pbase[c] (sum of n^2 before c)
// Query slot of pair t (t < 2 n^2) is 2*pbase + t; G slot is the same index.
// Sub-counters per coefficient block in the lookup histogram.
constexpr int kKeySub = 16;

// Green-pair reuse (reuse != 0): prev_verts holds the configuration's
// vertices of the plan's previous run (valid where prev_ok[c] is set), and
// g_up / g_dn still hold that run's Green values.  As the reference's pair
// cache (include/cdet/integrand_replica_randomly_sampled.h:57-86, :664-696),
// a vertex equal to a previous one (all four coordinates, bit for bit) takes
// its previous index, and a pair of two such vertices copies its previous
// value instead of being looked up: the value is the one the lookup computes,
// since every evaluation form gives a lookup the same result whatever its
// tile-mates.  cache[0] / cache[1] count reused and looked-up pairs.
// With rows_at_v1 (FP64 matrix-core plans) the diagonal pairs (i, i) are not
// looked up: they are lower-branch lookups at v = (2 tau) / tau - 1 = 1
// exactly, where T_b(1) = 1 for every b, so G = sum_a T_a(u) (sum_b c_ab), a
// Clenshaw over u of the block's row sums (ks.rowsum).  The rows of the
// configuration's diagonal pairs are staged in dynamic LDS (row_cap doubles,
// sized by the plan for all rows of its highest order, split into passes
// beyond kPlanRowCapMax) by coalesced workgroup loads, and one lane per pair
// runs the Clenshaw from there.
constexpr int kPlanPairsPerThread = (2 * kMaxOrder * kMaxOrder + 255) / 256;
constexpr int kPlanRowCapMax = 6144;  // doubles (48 KB)

__global__ __launch_bounds__(256) void k_plan(KScalars ks, const double* __restrict__ verts,
                                              const int32_t* __restrict__ cfg_order,
                                              const int64_t* __restrict__ cfg_vbase,
                                              const int64_t* __restrict__ cfg_pbase, double* __restrict__ g_up,
                                              double* __restrict__ g_dn, int32_t* __restrict__ q_key,
                                              int32_t* __restrict__ q_rank, double* __restrict__ q_u,
                                              double* __restrict__ q_v, int32_t* __restrict__ hist,
                                              int32_t* __restrict__ status, double* __restrict__ e_up,
                                              double* __restrict__ e_dn, double err_scale, int reuse,
                                              double* __restrict__ prev_verts, int32_t* __restrict__ prev_ok,
                                              unsigned long long* __restrict__ cache, int rows_at_v1, int row_cap) {
    __shared__ Vtx vs[2][kMaxOrder];
    __shared__ int pidx[kMaxOrder];
    __shared__ int s_hits, s_misses;
    __shared__ double s_du[2 * kMaxOrder];     // diagonal pair k = class n + i: u
    __shared__ int64_t s_dsp[2 * kMaxOrder];   // its sitepair
    __shared__ int s_dku[2 * kMaxOrder];       // its up-spin table kind
    __shared__ int s_dok[2 * kMaxOrder];       // evaluated from row sums in this run
    extern __shared__ double rows[];             // row_cap doubles
    const int c = blockIdx.x;
    const int n = cfg_order[c];
    if (threadIdx.x < 2 * kMaxOrder) s_dok[threadIdx.x] = 0;
    if (status && threadIdx.x == 0) status[c] = 0;
    const int sub = c & (kKeySub - 1);
    const int64_t vb = cfg_vbase[c];
    const int64_t qb = 2 * cfg_pbase[c];
    if (threadIdx.x < 2 * n) {
        const int cls = threadIdx.x / n;
        const int v = threadIdx.x - cls * n;
        vs[cls][v] = decode_vertex(verts + 4 * (vb + v), ks.beta, ks.N1, ks.N2, cls == 0 ? 2 : 1);
    }
    const bool have_prev = reuse && prev_ok[c] != 0;
    if (reuse && threadIdx.x < n) {
        int found = -1;
        if (have_prev) {
            const double* cur = verts + 4 * (vb + threadIdx.x);
            for (int k = 0; k < n && found < 0; ++k) {
                const double* pv = prev_verts + 4 * (vb + k);
                if (__double_as_longlong(cur[0]) == __double_as_longlong(pv[0]) &&
                    __double_as_longlong(cur[1]) == __double_as_longlong(pv[1]) &&
                    __double_as_longlong(cur[2]) == __double_as_longlong(pv[2]) &&
                    __double_as_longlong(cur[3]) == __double_as_longlong(pv[3]))
                    found = k;
            }
        }
        pidx[threadIdx.x] = found;
    }
    if (reuse && threadIdx.x == 0) {
        s_hits = 0;
        s_misses = 0;
    }
    __syncthreads();
    const int n2 = n * n;
    // previous values of the reusable pairs, read before any pair is written
    double keep_up[kPlanPairsPerThread], keep_dn[kPlanPairsPerThread];
    bool keep[kPlanPairsPerThread];
    if (reuse) {
#pragma unroll
        for (int r = 0; r < kPlanPairsPerThread; ++r) {
            const int t = threadIdx.x + r * 256;
            keep[r] = false;
            if (t < 2 * n2) {
                const int cls = t / n2;
                const int idx = t - cls * n2;
                const int i = idx / n;
                const int j = idx - i * n;
                if (pidx[i] >= 0 && pidx[j] >= 0) {
                    const int64_t src = qb + static_cast<int64_t>(cls) * n2 + pidx[i] * n + pidx[j];
                    keep[r] = true;
                    keep_up[r] = g_up[src];
                    keep_dn[r] = g_dn[src];
                }
            }
        }
        __syncthreads();
    }
    int hits = 0, misses = 0;
#pragma unroll
    for (int r = 0; r < kPlanPairsPerThread; ++r) {
        const int t = threadIdx.x + r * 256;
        if (t >= 2 * n2) break;
        const int cls = t / n2;
        const int idx = t - cls * n2;
        const int i = idx / n;
        const int j = idx - i * n;
        const Vtx& vi = vs[cls][i];
        const Vtx& vj = vs[cls][j];
        const bool upper = (i != j) && (vi.tau_unit > vj.tau_unit);
        const Desc d = make_desc(vi, vj, upper, ks.beta, ks.branch_eps, ks.N1, ks.N2);
        const int ku = cls == 0 ? CDET_TABLE_TWISTED_UP : CDET_TABLE_ORDINARY_UP;
        const int br = upper ? 0 : 1;
        const int64_t q = qb + t;
        if (reuse) {
            if (keep[r]) {
                g_up[q] = keep_up[r];
                g_dn[q] = keep_dn[r];
                q_key[q] = -1;
                ++hits;
                continue;
            }
            if (!d.corner) ++misses;
        }
        if (d.corner) {
            g_up[q] = ks.corner[ku][br][d.sitepair];
            g_dn[q] = ks.corner[ku + 1][br][d.sitepair];
            q_key[q] = -1;
            if (e_up) {  // corner values are read, not summed: no table-value error
                e_up[q] = 0.0;
                e_dn[q] = 0.0;
            }
            continue;
        }
        if (e_up) {  // error-bound kind: table-value error scale of the pair's block
            e_up[q] = err_scale * ks.abssum[ku][br][d.sitepair];
            e_dn[q] = err_scale * ks.abssum[ku + 1][br][d.sitepair];
        }
        if (rows_at_v1 && i == j && d.v == 1.0) {  // from the row sums below
            const int k = cls * n + i;
            s_du[k] = d.u;
            s_dsp[k] = d.sitepair;
            s_dku[k] = ku;
            s_dok[k] = 1;
            q_key[q] = -1;
            continue;
        }
        const int32_t key = static_cast<int32_t>((static_cast<int64_t>(cls) * 2 + br) * ks.SP + d.sitepair);
        const int32_t slot = key * kKeySub + sub;
        q_key[q] = slot;
        q_rank[q] = atomicAdd(&hist[slot], 1);
        q_u[q] = d.u;
        q_v[q] = d.v;
    }
    if (rows_at_v1) {
        __syncthreads();
        const int nu = ks.nu;
        const bool dn_rows = ks.rowsum[0] != ks.rowsum[1] || ks.rowsum[2] != ks.rowsum[3];
        const int per_pass = row_cap / (nu * (dn_rows ? 2 : 1));
        for (int k0 = 0; k0 < 2 * n; k0 += per_pass) {
            const int kc = min(per_pass, 2 * n - k0);
            for (int idx = threadIdx.x; idx < kc * nu; idx += blockDim.x) {
                const int kk = idx / nu;
                const int a = idx - kk * nu;
                const int k = k0 + kk;
                if (s_dok[k]) {
                    rows[kk * nu + a] = ks.rowsum[s_dku[k]][s_dsp[k] * nu + a];
                    if (dn_rows) rows[(per_pass + kk) * nu + a] = ks.rowsum[s_dku[k] + 1][s_dsp[k] * nu + a];
                }
            }
            __syncthreads();
            if (static_cast<int>(threadIdx.x) < kc) {
                const int k = k0 + threadIdx.x;
                if (s_dok[k]) {
                    const int cls = k >= n ? 1 : 0;
                    const int i = k - cls * n;
                    const int64_t q = qb + static_cast<int64_t>(cls) * n2 + i * n + i;
                    const double gu = clenshaw_row(rows + threadIdx.x * nu, nu - 1, s_du[k]);
                    g_up[q] = gu;
                    g_dn[q] = dn_rows ? clenshaw_row(rows + (per_pass + threadIdx.x) * nu, nu - 1, s_du[k]) : gu;
                }
            }
            __syncthreads();
        }
    }
    if (reuse) {
        // this run's vertices become the next run's previous ones (every
        // read of prev_verts above precedes the barrier)
        if (hits) atomicAdd(&s_hits, hits);
        if (misses) atomicAdd(&s_misses, misses);
        for (int k = threadIdx.x; k < 4 * n; k += blockDim.x) prev_verts[4 * vb + k] = verts[4 * vb + k];
        __syncthreads();
        if (threadIdx.x == 0) {
            prev_ok[c] = 1;
            if (cache) {
                atomicAdd(&cache[0], static_cast<unsigned long long>(s_hits));
                atomicAdd(&cache[1], static_cast<unsigned long long>(s_misses));
            }
        }
    }
}

// Table-evaluation planning: one thread per lookup (kind class, replicas,
// sites, tau1, tau2); upper branch iff tau1 > tau2.
__global__ __launch_bounds__(256) void k_plan_queries(KScalars ks, int64_t nq, const int32_t* __restrict__ qi,
                                                      const double* __restrict__ qt, double* __restrict__ g_up,
                                                      double* __restrict__ g_dn, int32_t* __restrict__ q_key,
                                                      int32_t* __restrict__ q_rank, double* __restrict__ q_u,
                                                      double* __restrict__ q_v, int32_t* __restrict__ hist) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int32_t* r = qi + 7 * q;
    const int cls = r[0] == 0 ? 0 : 1;
    Vtx vi, vj;
    vi.tau = qt[2 * q];
    vj.tau = qt[2 * q + 1];
    vi.tau_unit = vi.tau;
    vj.tau_unit = vj.tau;
    vi.replica = r[1];
    vj.replica = r[2];
    vi.x = r[3];
    vi.y = r[4];
    vj.x = r[5];
    vj.y = r[6];
    const bool upper = vi.tau > vj.tau;
    const Desc d = make_desc(vi, vj, upper, ks.beta, ks.branch_eps, ks.N1, ks.N2);
    const int ku = cls == 0 ? CDET_TABLE_TWISTED_UP : CDET_TABLE_ORDINARY_UP;
    const int br = upper ? 0 : 1;
    if (d.corner) {
        g_up[q] = ks.corner[ku][br][d.sitepair];
        g_dn[q] = ks.corner[ku + 1][br][d.sitepair];
        q_key[q] = -1;
        return;
    }
    const int32_t key = static_cast<int32_t>((static_cast<int64_t>(cls) * 2 + br) * ks.SP + d.sitepair);
    const int32_t slot = key * kKeySub + (blockIdx.x & (kKeySub - 1));
    q_key[q] = slot;
    q_rank[q] = atomicAdd(&hist[slot], 1);
    q_u[q] = d.u;
    q_v[q] = d.v;
}

__global__ void k_interleave(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                             double* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[2 * i] = a[i];
    out[2 * i + 1] = b[i];
}

// Exclusive scan of the per-block histogram and tile list, one workgroup.
// Offsets of the lookup groups and the tile list.  Each coefficient block
// (key) counts its lookups in kKeySub sub-counters (hist[key * kKeySub + sub])
// so that the atomics of the lookup planners spread over kKeySub addresses;
// the sub-groups of a key are laid out consecutively, so a key's lookups stay
// contiguous.  Tiles of <= tile lookups are cut per key.  The histogram is
// cleared after use (it starts cleared at allocation), and the per-XCD work
// counters are cleared for the contraction.
// Exclusive scan of (a, b) pairs over the workgroup: wave shuffles, then
// the wave totals.  Returns the workgroup totals.
__device__ inline int2 block_scan2(int a, int b, int& ea, int& eb, int2* s_w) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    int ia = a, ib = b;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int ta = __shfl_up(ia, off);
        const int tb = __shfl_up(ib, off);
        if (lane >= off) {
            ia += ta;
            ib += tb;
        }
    }
    if (lane == 63) s_w[w] = make_int2(ia, ib);
    __syncthreads();
    int pa = 0, pb = 0;
    int2 tot = make_int2(0, 0);
    const int nw = blockDim.x >> 6;
    for (int k = 0; k < nw; ++k) {
        const int2 v = s_w[k];
        if (k < w) {
            pa += v.x;
            pb += v.y;
        }
        tot.x += v.x;
        tot.y += v.y;
    }
    __syncthreads();
    ea = pa + ia - a;
    eb = pb + ib - b;
    return tot;
}

// Lookups of key k: the sum of its kKeySub sub-counters (read as int4).
__device__ inline int key_count(const int32_t* hist, int k, int4* hv) {
    const int4* src = reinterpret_cast<const int4*>(hist + static_cast<int64_t>(k) * kKeySub);
    int h = 0;
#pragma unroll
    for (int u = 0; u < kKeySub / 4; ++u) {
        hv[u] = src[u];
        h += hv[u].x + hv[u].y + hv[u].z + hv[u].w;
    }
    return h;
}

// Grouping, pass 1: per workgroup of kScanKeys keys, the total lookups and
// tiles (wg_tot[blockIdx]).
constexpr int kScanKeys = 256;
__global__ __launch_bounds__(kScanKeys) void k_key_totals(const int32_t* __restrict__ hist, int32_t nkeys,
                                                          int32_t tile, int2* __restrict__ wg_tot) {
    static_assert(kKeySub % 4 == 0, "sub-counters are read as int4");
    __shared__ int2 s_w[kScanKeys / 64];
    const int k = blockIdx.x * kScanKeys + threadIdx.x;
    int4 hv[kKeySub / 4];
    const int h = k < nkeys ? key_count(hist, k, hv) : 0;
    int eq, et;
    const int2 tot = block_scan2(h, (h + tile - 1) / tile, eq, et, s_w);
    if (threadIdx.x == 0) wg_tot[blockIdx.x] = tot;
}

// Grouping, pass 2: offsets of the lookup groups and the tile list.  Each
// coefficient block (key) counts its lookups in kKeySub sub-counters
// (hist[key * kKeySub + sub]) so that the atomics of the lookup planners
// spread over kKeySub addresses; the sub-groups of a key are laid out
// consecutively, so a key's lookups stay contiguous.  Tiles of <= tile
// lookups are cut per key.  The histogram is cleared after use (it starts
// cleared at allocation).
__global__ __launch_bounds__(kScanKeys) void k_scan_keys(int32_t* __restrict__ hist, int32_t nkeys, int32_t tile,
                                                         const int2* __restrict__ wg_tot,
                                                         int32_t* __restrict__ key_off, int32_t* __restrict__ tile_key,
                                                         int32_t* __restrict__ tile_start,
                                                         int32_t* __restrict__ tile_len, int32_t* __restrict__ n_tiles) {
    __shared__ int2 s_w[kScanKeys / 64];
    __shared__ int2 s_base;
    const int tid = threadIdx.x;
    // offsets of this workgroup: totals of the workgroups before it
    int bq = 0, bt = 0;
    for (int w = tid; w < static_cast<int>(blockIdx.x); w += kScanKeys) {
        const int2 v = wg_tot[w];
        bq += v.x;
        bt += v.y;
    }
    int eq0, et0;
    const int2 pre = block_scan2(bq, bt, eq0, et0, s_w);
    if (tid == 0) s_base = pre;
    __syncthreads();
    const int cq = s_base.x, ct = s_base.y;
    const int k = blockIdx.x * kScanKeys + tid;
    int4 hv[kKeySub / 4];
    const int h = k < nkeys ? key_count(hist, k, hv) : 0;
    int eq, et;
    const int2 tot = block_scan2(h, (h + tile - 1) / tile, eq, et, s_w);
    if (k < nkeys) {
        int off = cq + eq;
        int4* dst = reinterpret_cast<int4*>(key_off + static_cast<int64_t>(k) * kKeySub);
        int4* hz = reinterpret_cast<int4*>(hist + static_cast<int64_t>(k) * kKeySub);
#pragma unroll
        for (int u = 0; u < kKeySub / 4; ++u) {
            int4 o;
            o.x = off;
            off += hv[u].x;
            o.y = off;
            off += hv[u].y;
            o.z = off;
            off += hv[u].z;
            o.w = off;
            off += hv[u].w;
            dst[u] = o;
            hz[u] = make_int4(0, 0, 0, 0);
        }
        int ot = ct + et;
        for (int s = 0; s < h; s += tile) {
            tile_key[ot] = k;
            tile_start[ot] = cq + eq + s;
            tile_len[ot] = min(tile, h - s);
            ++ot;
        }
    }
    if (blockIdx.x == gridDim.x - 1 && tid == 0) *n_tiles = ct + tot.y;
}

__global__ __launch_bounds__(256) void k_scatter(int64_t nq, const int32_t* __restrict__ q_key,
                                                 const int32_t* __restrict__ q_rank, const double* __restrict__ q_u,
                                                 const double* __restrict__ q_v, const int32_t* __restrict__ key_off,
                                                 double* __restrict__ s_u, double* __restrict__ s_v,
                                                 int32_t* __restrict__ s_dest) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int32_t key = q_key[q];
    if (key < 0) return;
    const int32_t pos = key_off[key] + q_rank[q];
    s_u[pos] = q_u[q];
    s_v[pos] = q_v[q];
    s_dest[pos] = static_cast<int32_t>(q);
}

// Row Clenshaw of ROWS consecutive rows (a, a-1, ...) in lockstep: ROWS
// independent dependency chains per lane, each the reference's recurrence
// (include/cdet/g_funcs.h:450-461) for one row.
template <int ROWS>
__device__ inline void clenshaw_rows(const double* __restrict__ c, int a, int nv, double v, double two_v,
                                     double* __restrict__ r) {
    double b1[ROWS], b2[ROWS];
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
        b1[q] = 0.0;
        b2[q] = 0.0;
    }
    for (int k = nv - 1; k >= 1; --k) {
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
            const double bk = c[static_cast<int64_t>(a - q) * nv + k] + two_v * b1[q] - b2[q];
            b2[q] = b1[q];
            b1[q] = bk;
        }
    }
#pragma unroll
    for (int q = 0; q < ROWS; ++q) r[q] = c[static_cast<int64_t>(a - q) * nv] + b1[q] * v - b2[q];
}

typedef double d8v __attribute__((ext_vector_type(8)));

// clenshaw_rows for nv % 8 == 0: each row's coefficients are fetched eight at
// a time (one 64-byte wave-uniform load, i.e. one scalar-cache line) and the
// eight recurrence steps of the chunk run from registers.
template <int ROWS>
__device__ inline void clenshaw_rows8(const double* __restrict__ c, int a, int nv, double v, double two_v,
                                      double* __restrict__ r) {
    double b1[ROWS], b2[ROWS], c0[ROWS];
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
        b1[q] = 0.0;
        b2[q] = 0.0;
    }
    for (int kc = nv - 8; kc >= 0; kc -= 8) {
        d8v x[ROWS];
#pragma unroll
        for (int q = 0; q < ROWS; ++q) x[q] = *reinterpret_cast<const d8v*>(c + static_cast<int64_t>(a - q) * nv + kc);
        if (kc > 0) {
#pragma unroll
            for (int j = 7; j >= 0; --j) {
#pragma unroll
                for (int q = 0; q < ROWS; ++q) {
                    const double bk = x[q][j] + two_v * b1[q] - b2[q];
                    b2[q] = b1[q];
                    b1[q] = bk;
                }
            }
        } else {
#pragma unroll
            for (int j = 7; j >= 1; --j) {
#pragma unroll
                for (int q = 0; q < ROWS; ++q) {
                    const double bk = x[q][j] + two_v * b1[q] - b2[q];
                    b2[q] = b1[q];
                    b1[q] = bk;
                }
            }
#pragma unroll
            for (int q = 0; q < ROWS; ++q) c0[q] = x[q][0];
        }
    }
#pragma unroll
    for (int q = 0; q < ROWS; ++q) r[q] = c0[q] + b1[q] * v - b2[q];
}

// G for one query from one coefficient block, rows produced four at a time
// from a = nu-1 downwards and consumed in that order by the Clenshaw over u;
// bit-identical to chebyshev_sum_2d_pair_from_ptrs (include/cdet/g_funcs.h:687-763).
__device__ inline double cheb2d_exact(const double* __restrict__ c, int nu, int nv, double u, double v) {
    UClenshaw fu;
    fu.init(u, nu);
    double g = 0.0;
    const int dv = nv - 1;
    if (dv < 2) {
        for (int a = nu - 1; a >= 0; --a) g = fu.feed(a, clenshaw_row(c + static_cast<int64_t>(a) * nv, dv, v));
        return g;
    }
    const double two_v = 2.0 * v;
    int a = nu - 1;
    if ((nv & 7) == 0) {
        for (; a >= 3; a -= 4) {
            double r[4];
            clenshaw_rows8<4>(c, a, nv, v, two_v, r);
            g = fu.feed(a, r[0]);
            g = fu.feed(a - 1, r[1]);
            g = fu.feed(a - 2, r[2]);
            g = fu.feed(a - 3, r[3]);
        }
        for (; a >= 0; --a) {
            double r[1];
            clenshaw_rows8<1>(c, a, nv, v, two_v, r);
            g = fu.feed(a, r[0]);
        }
        return g;
    }
    for (; a >= 3; a -= 4) {
        double r[4];
        clenshaw_rows<4>(c, a, nv, v, two_v, r);
        g = fu.feed(a, r[0]);
        g = fu.feed(a - 1, r[1]);
        g = fu.feed(a - 2, r[2]);
        g = fu.feed(a - 3, r[3]);
    }
    for (; a >= 0; --a) {
        double r[1];
        clenshaw_rows<1>(c, a, nv, v, two_v, r);
        g = fu.feed(a, r[0]);
    }
    return g;
}

// Exact batched contraction: one wave per tile of <= 64 queries sharing a
// coefficient block; lane = query.  The block address is wave-uniform, so the
// coefficients stream through the scalar cache and are broadcast to all
// lanes.  Tiles are split evenly over the XCDs in contiguous ranges.
__global__ __launch_bounds__(256) void k_contract_exact(KScalars ks, const int32_t* __restrict__ n_tiles,
                                                        const int32_t* __restrict__ tile_key,
                                                        const int32_t* __restrict__ tile_start,
                                                        const int32_t* __restrict__ tile_len,
                                                        const double* __restrict__ s_u, const double* __restrict__ s_v,
                                                        const int32_t* __restrict__ s_dest, double* __restrict__ g_up,
                                                        double* __restrict__ g_dn) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int waves_per_block = blockDim.x >> 6;
    const TileRange tr = xcd_tiles(*n_tiles, wave, waves_per_block);
    for (int t = tr.first; t < tr.end; t += tr.stride) {
        const int tile = __builtin_amdgcn_readfirstlane(t);
        const int32_t key = __builtin_amdgcn_readfirstlane(tile_key[tile]);
        const int32_t start = __builtin_amdgcn_readfirstlane(tile_start[tile]);
        const int32_t len = __builtin_amdgcn_readfirstlane(tile_len[tile]);
        const double* cu;
        const double* cd;
        bool same;
        ks.key_tables(key, cu, cd, same);
        const bool active = lane < len;
        const int pos = start + (active ? lane : 0);
        const double u = s_u[pos];
        const double v = s_v[pos];
        const double gu = cheb2d_exact(cu, ks.nu, ks.nv, u, v);
        const double gd = same ? gu : cheb2d_exact(cd, ks.nu, ks.nv, u, v);
        if (active) {
            const int32_t dest = s_dest[pos];
            g_up[dest] = gu;
            g_dn[dest] = gd;
        }
    }
}

// Order-bucketed principal minors and connected values, two kernels per
// order bucket (list index li = position in the bucket's configuration list):
//
// k_fpm   one workgroup per spin matrix (class 0 = twisted, rc = 2; 1 =
//         ordinary, rc = 1).  Builds the vertex matrix
//         (include/cdet/integrand_replica_randomly_sampled.h:751-761), runs
//         the Schur-complement tree and writes the principal minors to
//         pm[(4 li + 2 class + spin) 2^n + mask].  A matrix whose pivot needs the reference's
//         fallback is queued instead (entry = (li << 3) | (class << 2) | spin
//         mask) and k_fallback_dets writes its pivoted determinants to the same
//         place.  Identical up and down tables (same_spin) use spin 0 for both.
// k_cdet  per class: a = pm_up * pm_down, subset recursion, assembly; one
//         workgroup per (li, class), w = 2 li + class, or (paired) one per
//         configuration li taking the twisted class, then the ordinary one.
//   mode 0 (all subsets): out[obase, obase + 2^n) receives twisted -
//          2 * ordinary: formed in place when paired, else the ordinary
//          values go to out_ord and k_combine_all forms it.
//   mode 1 (single order): the class value as a double-double pair into
//          single_tmp[4 c + 2 class + {0, 1}]; k_combine_single forms the
//          double-double difference.
//   mode 2: conditioning scales of the mode-0 values (twisted + 2 *
//          ordinary), same layout.
// With GSCR (orders whose workspace exceeds LDS) the workspace is global
// scratch and a fixed grid loops over the work items; GSCR is a template
// argument so that LDS accesses compile to ds_read/ds_write.
struct FpmLds {
    static __host__ __device__ inline size_t doubles(int n) {
        const size_t sc = size_t(1) << n;
        const size_t inv = n >= 2 ? (size_t(1) << (n - 2)) : 1;
        return static_cast<size_t>(n) * n + sc + 2 * static_cast<size_t>(fpm_list_size(n)) + inv + 64;
    }
    static __host__ __device__ inline size_t bytes(int n) { return sizeof(double) * doubles(n); }
};
struct CdetLds {
    static __host__ __device__ inline size_t doubles(int n, int mode) {
        const size_t sc = size_t(1) << n;
        return (mode == 2 ? 3 : 2) * sc + 2 * static_cast<size_t>(n) + sc / 4 + 1;
    }
    static __host__ __device__ inline size_t bytes(int n, int mode) { return sizeof(double) * doubles(n, mode); }
};

template <bool GSCR, int NT>
__global__ __launch_bounds__(NT) void k_fpm(KScalars ks, int n, int items, const int32_t* __restrict__ cfg_list,
                                             const int64_t* __restrict__ cfg_pbase, const double* __restrict__ g_up,
                                             const double* __restrict__ g_dn, double* __restrict__ pm,
                                             int32_t* __restrict__ fb_list, int32_t* __restrict__ fb_count,
                                             double* __restrict__ gscratch, size_t gscratch_stride) {
    extern __shared__ double lds_dyn[];
    __shared__ int s_flag;
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const uint32_t sc = 1u << n;
    const int n2 = n * n;
    const int list = fpm_list_size(n);
    double* A = GSCR ? gscratch + static_cast<size_t>(blockIdx.x) * gscratch_stride : lds_dyn;
    double* fo = A + n2;
    double* bufA = fo + sc;
    double* bufB = bufA + list;
    double* inv = bufB + list;
    double* red = inv + (n >= 2 ? (1 << (n - 2)) : 1);
    // Work items: 2 (li, class) pairs x 2 spins, or x 1 when both classes have
    // identical spin tables.
    const bool both_same = ks.same_spin[0] && ks.same_spin[1];
    for (int w = blockIdx.x; w < items; w += gridDim.x) {
        const int li = both_same ? (w >> 1) : (w >> 2);
        const int cls = both_same ? (w & 1) : ((w >> 1) & 1);
        const int spin = both_same ? 0 : (w & 1);
        const bool same = ks.same_spin[cls] != 0;
        if (spin == 1 && same) continue;  // uniform across the workgroup
        const int c = cfg_list[li];
        const double* g = (spin ? g_dn : g_up) + 2 * cfg_pbase[c] + static_cast<int64_t>(cls) * n2;
        const double shift = spin ? ks.ddn : ks.dup;
        for (int idx = tid; idx < n2; idx += nt) {
            const int i = idx / n;
            double v = g[idx];
            if (i * n + i == idx) v += shift;
            A[idx] = v * ks.scale;
        }
        __syncthreads();
        if (fpm_block<false>(A, n, ks.tol, ks.shift, fo, bufA, bufB, inv, red, &s_flag)) {
            if (tid == 0) {
                const int slot = atomicAdd(fb_count, 1);
                fb_list[slot] = (li << 3) | (cls << 2) | (same ? 3 : (1 << spin));
            }
        } else {
            double* dst = pm + static_cast<size_t>(4 * li + 2 * cls + spin) * sc;
            for (uint32_t m = tid; m < sc; m += nt) dst[m] = fo[m];
        }
        __syncthreads();
    }
}

// Pivoted determinants of every principal submatrix of the queued fallback
// spin matrices (include/cdet/fast_principle_minors.h:39-115, :180-186),
// spread over a fixed grid: work item = (queue entry, spin, chunk of
// det_rows_item<NM>() masks), one mask per 16-lane group with one row of the
// working matrix per lane (det_minor_rows), four masks per wave and pass.
// NM = 10 serves orders up to 10 (90 VGPRs), NM = 14 the orders above.
// Counts the fallbacks (include/cdet/runtime_counters.h:27) and flags the
// configuration's status.  With cfg_order non-null configuration c has order
// cfg_order[c] <= n (n sizes the chunks and the LDS); minors go to pm[(4 li +
// 2 class + spin) 2^pm_log + mask].
// (Also measured: the whole k x k matrix in one thread's registers, 256
// VGPRs: the same C2 rates alone and with four batches in flight; working
// matrices interleaved in LDS, one mask per thread: the order-10 minors stage
// 94 -> 120 us, and 432 us for an order-14 fallback matrix of the C5 batch.)
constexpr int kDetThreads = 64;
constexpr int kDetRegMax = 10;  // orders served by the NM = 10 form
// masks per work item: one pass of four masks up to order 10, four passes
// above (their matrices take longer to stage than to use once)
template <int NM>
constexpr int det_rows_item() {
    return NM <= kDetRegMax ? 4 : 16;
}
template <int NM>
__global__ __launch_bounds__(kDetThreads) void k_fallback_dets(
    KScalars ks, int n_max, const int32_t* __restrict__ cfg_list, const int64_t* __restrict__ cfg_pbase,
    const double* __restrict__ g_up, const double* __restrict__ g_dn, const int32_t* __restrict__ fb_list,
    const int32_t* __restrict__ fb_count, double* __restrict__ pm, int32_t* __restrict__ status,
    uint64_t* __restrict__ fallback_total, const int32_t* __restrict__ cfg_order, int pm_log) {
    extern __shared__ double lds_dyn[];
    const int tid = threadIdx.x;
    constexpr int kItem = det_rows_item<NM>();
    const int chunks = static_cast<int>(((1u << n_max) + kItem - 1) / kItem);
    double* A = lds_dyn;
    const int total = *fb_count * 2 * chunks;
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        const int e = w / (2 * chunks);
        const int spin = (w / chunks) & 1;
        const int chunk = w - (w / chunks) * chunks;
        const int entry = fb_list[e];
        const int li = entry >> 3;
        const int cls = (entry >> 2) & 1;
        const int spins = entry & 3;
        // spin 1 of identical spin tables is never read (the up minors serve both)
        if (!((spins >> spin) & 1) || (spin == 1 && ks.same_spin[cls])) continue;
        const int c = cfg_list[li];
        const int n = cfg_order ? cfg_order[c] : n_max;
        const int n2 = n * n;
        const uint32_t sc = 1u << n;
        if (chunk == 0 && tid == 0 && spin == __builtin_ctz(spins)) {
            atomicAdd(reinterpret_cast<unsigned long long*>(fallback_total),
                      static_cast<unsigned long long>(__builtin_popcount(spins)));
            if (status) atomicOr(&status[c], 2);
        }
        const double* g = (spin == 0 ? g_up : g_dn) + 2 * cfg_pbase[c] + static_cast<int64_t>(cls) * n2;
        const double shift = spin == 0 ? ks.dup : ks.ddn;
        __syncthreads();
        for (int idx = tid; idx < n2; idx += kDetThreads) {
            const int i = idx / n;
            double v = g[idx];
            if (i * n + i == idx) v += shift;
            A[idx] = v * ks.scale;
        }
        __syncthreads();
        double* dst = pm + (static_cast<size_t>(4 * li + 2 * cls + spin) << pm_log);
        // one mask per 16-lane group, every lane of the wave in the call
        const int r = tid & 15;
        const int gbase = (tid & 63) - r;
#pragma unroll
        for (int p = 0; p < kItem; p += kDetThreads / 16) {
            const uint32_t m = static_cast<uint32_t>(chunk * kItem + p + (tid >> 4));
            const double det = det_minor_rows<NM>(A, n, m < sc ? m : 0u, r, gbase);
            if (m < sc && r == 0) dst[m] = det;  // mask 0: 1
        }
    }
}

// masks_by_pop / pop_off: the tables of all orders (order k at k 2^kMaxOrder
// and k (kMaxOrder + 2)).  With cfg_order non-null configuration c has order
// cfg_order[c] <= n (n sized the LDS or gscratch_stride); the minors of list
// entry li are at pm + (4 li + 2 class + spin) 2^pm_log.
// LAY: where the work item's arrays live.
//   kCdetLayLds    a, dp (rhs), dU, dD and the mask list in LDS;
//   kCdetLayGlobal all of them in global scratch (gscratch_stride doubles per
//                  workgroup): orders whose arrays exceed the LDS;
//   kCdetLaySplit  a (read only once formed) in global scratch, dp, dU and dD
//                  in LDS, the mask list read from masks_all: order 14 in
//                  modes 0 and 1, whose 2^14 connected values (128 KB) fit
//                  the LDS where a and dp together do not.
constexpr int kCdetLayLds = 0;
constexpr int kCdetLayGlobal = 1;
constexpr int kCdetLaySplit = 2;
struct CdetSplitLds {
    static __host__ __device__ inline size_t doubles(int n) {
        return (size_t(1) << n) + 2 * static_cast<size_t>(n);
    }
    static __host__ __device__ inline size_t bytes(int n) { return sizeof(double) * doubles(n); }
};
template <int LAY, int NT>
__global__ __launch_bounds__(NT) void k_cdet(KScalars ks, int n_max, int mode, int items,
                                              const int32_t* __restrict__ cfg_list,
                                              const int64_t* __restrict__ cfg_pbase,
                                              const int64_t* __restrict__ cfg_obase, const double* __restrict__ g_up,
                                              const double* __restrict__ g_dn, const double* __restrict__ pm,
                                              double* __restrict__ out, double* __restrict__ out_ord,
                                              double* __restrict__ single_tmp, int32_t* __restrict__ status,
                                              const uint16_t* __restrict__ masks_all,
                                              const int* __restrict__ pop_off_all, double* __restrict__ gscratch,
                                              size_t gscratch_stride, int32_t* __restrict__ fb_count, int paired,
                                              const int32_t* __restrict__ cfg_order, int pm_log) {
    extern __shared__ double lds_dyn[];
    __shared__ __attribute__((aligned(16))) double tb[kFoldWaves * kFoldSteps * 64];
    __shared__ int s_popoff[kMaxOrder + 2];  // popcount-layer offsets, read every layer
    // The bucket's fallback queue has been consumed (k_fallback_dets ran
    // before); clear its length for the next run.
    if (fb_count && blockIdx.x == 0 && threadIdx.x == 0) *fb_count = 0;  // wave-wide chain terms of the subset recursion
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    constexpr bool GSCR = LAY != kCdetLayLds;
    constexpr bool SPLIT = LAY == kCdetLaySplit;
    double* a = GSCR ? gscratch + static_cast<size_t>(blockIdx.x) * gscratch_stride : lds_dyn;
    // paired (modes 0 and 2): one work item per configuration, its two
    // classes in turn, the ordinary class combined into the twisted values in
    // place; otherwise one work item per (configuration, class)
    const int classes = paired ? 2 : 1;
    for (int w = blockIdx.x; w < items; w += gridDim.x) {
      for (int ci = 0; ci < classes; ++ci) {
        const int li = paired ? w : (w >> 1);
        const int cls = paired ? ci : (w & 1);
        const int c = cfg_list[li];
        const int n = cfg_order ? cfg_order[c] : n_max;
        const uint32_t sc = 1u << n;
        const int n2 = n * n;
        double* dp = SPLIT ? lds_dyn : a + sc;
        double* rhs = dp + sc;  // mode 2 only (never split)
        double* dU = rhs + (mode == 2 ? sc : 0);
        double* dD = dU + n;
        const uint16_t* masks_by_pop = masks_all + (static_cast<size_t>(n) << kMaxOrder);
        // the mask list of the recursion: a copy in the work item's arrays, or
        // the table itself when split
        uint16_t* lmask = SPLIT ? nullptr : reinterpret_cast<uint16_t*>(dD + n);
        const uint16_t* lm = SPLIT ? masks_by_pop : lmask;
        for (int k = tid; k < kMaxOrder + 2; k += nt) s_popoff[k] = pop_off_all[n * (kMaxOrder + 2) + k];
        const bool same = ks.same_spin[cls] != 0;
        const int64_t pb = 2 * cfg_pbase[c] + static_cast<int64_t>(cls) * n2;
        // unscaled shifted diagonal entries for the singletons (:751-758)
        for (int i = tid; i < n; i += nt) {
            dU[i] = g_up[pb + i * n + i] + ks.dup;
            dD[i] = g_dn[pb + i * n + i] + ks.ddn;
        }
        if (n == 1) {  // order-1 single value (:173-213), no minors
            __syncthreads();
            if (tid == 0) {
                single_tmp[4 * c + 2 * cls] = ks.c1[cls] * (dU[0] * dD[0] - ks.counter);
                single_tmp[4 * c + 2 * cls + 1] = 0.0;
            }
            __syncthreads();
            continue;
        }
        const double* pu = pm + (static_cast<size_t>(4 * li + 2 * cls) << pm_log);
        const double* pd = same ? pu : pu + (size_t(1) << pm_log);
        // a[mask] = pm_up * pm_down with a[0] = 0; the global loads of up to
        // four masks per thread are issued before any product is stored
        {
            constexpr int kPre = 4;
            double vu[kPre], vd[kPre];
            uint16_t mk[kPre];
            for (uint32_t m0 = tid; m0 < sc; m0 += kPre * nt) {
#pragma unroll
                for (int q = 0; q < kPre; ++q) {
                    const uint32_t m = m0 + q * nt;
                    if (m < sc) {
                        vu[q] = pu[m];
                        vd[q] = pd[m];
                        if (!SPLIT) mk[q] = masks_by_pop[m];
                    }
                }
#pragma unroll
                for (int q = 0; q < kPre; ++q) {
                    const uint32_t m = m0 + q * nt;
                    if (m < sc) {
                        a[m] = (m == 0) ? 0.0 : vu[q] * vd[q];
                        if (!SPLIT) lmask[m] = mk[q];
                    }
                }
            }
        }
        __syncthreads();
        if (mode == 0 || mode == 2) {
            // mode 2 writes the magnitude scale of each value instead of the value:
            // |prefactor| * (|a[mask]| + sum |terms|), |value| for singletons.
            ztransform_layers(a, dp, n, ks.thr, ks.filter != 0, lm, s_popoff, mode == 2 ? rhs : nullptr, false, tb,
                              ks.group16_R[n], ks.group64_R[n], ks.tree_row(n));
            // Assembly (:833-877).
            // paired: twisted - 2 * ordinary (include/cdet/cdet_integrand_evaluator.h:348-350),
            // or for magnitude scales twisted + 2 * ordinary, formed in place;
            // each thread reads back only the twisted values it wrote itself.
            // Otherwise the ordinary values go to out_ord for k_combine_all.
            double* dst = (cls == 0 || paired ? out : out_ord) + cfg_obase[c];
            bool bad = false;
            for (uint32_t m = tid; m < sc; m += nt) {
                double v;
                if (m == 0) {
                    v = 0.0;
                } else if ((m & (m - 1)) == 0) {
                    const int i = __builtin_ctz(m);
                    v = ks.c1[cls] * (dU[i] * dD[i] - ks.counter);
                    if (mode == 2) v = fabs(v);
                } else {
                    v = mode == 2 ? fabs(ks.pref[cls][__builtin_popcount(m)]) * rhs[m]
                                  : ks.pref[cls][__builtin_popcount(m)] * dp[m];
                }
                if (!isfinite(v)) bad = true;
                dst[m] = (cls == 0 || !paired) ? v : (mode == 2 ? dst[m] + 2.0 * v : dst[m] - 2.0 * v);
            }
            if (bad && status) atomicOr(&status[c], 1);
        } else {
            ztransform_layers(a, dp, n, ks.thr, ks.filter != 0, lm, s_popoff, nullptr, true, tb, ks.group16_R[n],
                              ks.group64_R[n], ks.tree_row(n));
            const double cres = dp[sc - 1];
            if (tid == 0) {
                const DD v = dd_mul(DD{ks.pref_hi[cls][n], ks.pref_lo[cls][n]}, DD{cres, 0.0});
                single_tmp[4 * c + 2 * cls] = v.hi;
                single_tmp[4 * c + 2 * cls + 1] = v.lo;
                if (!isfinite(v.hi + v.lo) && status) atomicOr(&status[c], 1);
            }
        }
        __syncthreads();
      }
    }
}

// Paired classes in one pass (orders up to kPairOrderMax): one workgroup per
// configuration runs the subset recursions of the twisted and the ordinary
// class layer by layer together (ztransform_layers_pair) and forms
// twisted - 2 * ordinary (mode 0) or the magnitude scales twisted + 2 *
// ordinary (mode 2) from LDS: the same values, bit for bit, as k_cdet's paired
// form, which runs the classes one after the other and reads the twisted
// values back from HBM.
constexpr int kPairOrderMax = 10;
struct CdetPairLds {
    static __host__ __device__ inline size_t doubles(int n, int mode) {
        const size_t sc = size_t(1) << n;
        return (mode == 2 ? 6 : 4) * sc + 4 * static_cast<size_t>(n) + sc / 4 + 1;
    }
    static __host__ __device__ inline size_t bytes(int n, int mode) { return sizeof(double) * doubles(n, mode); }
};

// Recursion and assembly of one configuration's two classes once a[class][mask]
// = pm_up * pm_down is in LDS (a[class][0] = 0) together with the unscaled
// shifted diagonals dU, dD (class 1 at offset n): the connected values
// (include/cdet/integrand_replica_randomly_sampled.h:292-347), the singletons
// and prefactors (:833-877) and twisted - 2 * ordinary
// (include/cdet/cdet_integrand_evaluator.h:348-350) into dst; mode 2 writes
// the magnitude scales twisted + 2 * ordinary instead.  Non-finite values set
// status bit 0.  Ends with a workgroup barrier.
__device__ __forceinline__ void pair_finish(const KScalars& ks, int n, int mode, const double* dU, const double* dD,
                                            double* a, double* dp, double* rhs, const uint16_t* lmask,
                                            const int* s_popoff, double* tb, double* __restrict__ dst,
                                            int32_t* status_c) {
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const uint32_t sc = 1u << n;
    ztransform_layers_pair(a, dp, n, ks.thr, ks.filter != 0, lmask, s_popoff, mode == 2 ? rhs : nullptr, tb,
                           ks.group16_R[n], ks.group64_R[n], ks.tree_row(n));
    bool bad = false;
    for (uint32_t m = tid; m < sc; m += nt) {
        double v[2];
#pragma unroll
        for (int cls = 0; cls < 2; ++cls) {
            if (m == 0) {
                v[cls] = 0.0;
            } else if ((m & (m - 1)) == 0) {
                const int i = __builtin_ctz(m);
                v[cls] = ks.c1[cls] * (dU[cls * n + i] * dD[cls * n + i] - ks.counter);
                if (mode == 2) v[cls] = fabs(v[cls]);
            } else {
                const int k = __builtin_popcount(m);
                v[cls] = mode == 2 ? fabs(ks.pref[cls][k]) * rhs[cls * sc + m] : ks.pref[cls][k] * dp[cls * sc + m];
            }
            if (!isfinite(v[cls])) bad = true;
        }
        dst[m] = mode == 2 ? v[0] + 2.0 * v[1] : v[0] - 2.0 * v[1];
    }
    if (bad && status_c) atomicOr(status_c, 1);
    __syncthreads();
}

// Unscaled shifted diagonal entries of both classes
// (include/cdet/integrand_replica_randomly_sampled.h:751-758), thread q0
// onward in steps of stride.
__device__ __forceinline__ void pair_diagonals(const KScalars& ks, int n, int64_t pb, const double* __restrict__ g_up,
                                               const double* __restrict__ g_dn, double* dU, double* dD, int q0,
                                               int stride) {
    const int n2 = n * n;
    for (int q = q0; q < 2 * n; q += stride) {
        const int cls = q >= n ? 1 : 0;
        const int i = q - cls * n;
        const int64_t at = pb + static_cast<int64_t>(cls) * n2 + i * n + i;
        dU[q] = g_up[at] + ks.dup;
        dD[q] = g_dn[at] + ks.ddn;
    }
}

// With defer_count non-null the work items are the defer_count[0]
// configurations of defer_list (the fused kernel's deferred ones, below)
// instead of cfg_list[0 .. items); the grid's last workgroup to finish
// resets both queue words (defer_count[1] counts finished workgroups).
// masks_all / pop_off_all hold the tables of every order; with cfg_order
// non-null configuration c has order cfg_order[c] <= n (n sized the LDS) and
// the minors of list entry li are at pm + (4 li + 2 class) 2^pm_log.
__global__ __launch_bounds__(256) void k_cdet_pair(KScalars ks, int n_max, int mode, int items,
                                                   const int32_t* __restrict__ cfg_list,
                                                   const int64_t* __restrict__ cfg_pbase,
                                                   const int64_t* __restrict__ cfg_obase,
                                                   const double* __restrict__ g_up, const double* __restrict__ g_dn,
                                                   const double* __restrict__ pm, double* __restrict__ out,
                                                   int32_t* __restrict__ status,
                                                   const uint16_t* __restrict__ masks_all,
                                                   const int* __restrict__ pop_off_all, int32_t* __restrict__ fb_count,
                                                   const int32_t* __restrict__ defer_list,
                                                   int32_t* __restrict__ defer_count,
                                                   const int32_t* __restrict__ cfg_order, int pm_log) {
    extern __shared__ double lds_dyn[];
    __shared__ __attribute__((aligned(16))) double tb[kFoldWaves * kFoldSteps * 64];
    __shared__ int s_popoff[kMaxOrder + 2];
    if (fb_count && blockIdx.x == 0 && threadIdx.x == 0) *fb_count = 0;  // the bucket's fallback queue is consumed
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    int n = -1;  // order of the tables in LDS
    if (defer_count) items = __atomic_load_n(defer_count, __ATOMIC_RELAXED);
    for (int w = blockIdx.x; w < items; w += gridDim.x) {
        const int li = defer_count ? defer_list[w] : w;
        const int c = cfg_list[li];
        const int nc = cfg_order ? cfg_order[c] : n_max;
        const uint32_t sc = 1u << nc;
        double* a = lds_dyn;
        double* dp = a + 2 * sc;
        double* rhs = dp + 2 * sc;  // mode 2 only
        double* dU = rhs + (mode == 2 ? 2 * sc : 0);
        double* dD = dU + 2 * nc;
        uint16_t* lmask = reinterpret_cast<uint16_t*>(dD + 2 * nc);
        if (nc != n) {
            // a change of order moves lmask: every read of the previous
            // configuration's tables ended at pair_finish's closing barrier
            n = nc;
            for (int k = tid; k < kMaxOrder + 2; k += nt) s_popoff[k] = pop_off_all[n * (kMaxOrder + 2) + k];
            const uint16_t* masks_by_pop = masks_all + (static_cast<size_t>(n) << kMaxOrder);
            for (uint32_t m = tid; m < sc; m += nt) lmask[m] = masks_by_pop[m];
        }
        pair_diagonals(ks, n, 2 * cfg_pbase[c], g_up, g_dn, dU, dD, tid, nt);
        // a[class][mask] = pm_up * pm_down, a[class][0] = 0; four masks per
        // thread in flight
        {
            constexpr int kPre = 4;
            double vu[kPre], vd[kPre];
            for (uint32_t m0 = tid; m0 < 2 * sc; m0 += kPre * nt) {
#pragma unroll
                for (int q = 0; q < kPre; ++q) {
                    const uint32_t m = m0 + q * nt;
                    if (m < 2 * sc) {
                        const int cls = m >= sc ? 1 : 0;
                        const uint32_t mm = m - cls * sc;
                        const double* pu = pm + (static_cast<size_t>(4 * li + 2 * cls) << pm_log);
                        const double* pd = ks.same_spin[cls] ? pu : pu + (size_t(1) << pm_log);
                        vu[q] = pu[mm];
                        vd[q] = pd[mm];
                    }
                }
#pragma unroll
                for (int q = 0; q < kPre; ++q) {
                    const uint32_t m = m0 + q * nt;
                    if (m < 2 * sc) a[m] = (m & (sc - 1)) == 0 ? 0.0 : vu[q] * vd[q];
                }
            }
        }
        __syncthreads();
        pair_finish(ks, n, mode, dU, dD, a, dp, rhs, lmask, s_popoff, tb, out + cfg_obase[c],
                    status ? status + c : nullptr);
    }
    if (defer_count) {
        __syncthreads();
        if (tid == 0) {
            __threadfence();
            if (atomicAdd(defer_count + 1, 1) == static_cast<int>(gridDim.x) - 1) {
                defer_count[0] = 0;
                defer_count[1] = 0;
            }
        }
    }
}

// Principal minors and paired subset recursion of one configuration in one
// workgroup (orders 2 .. kPairOrderMax, identical up and down tables in both
// classes, modes 0 and 2).  Waves 0 and 1 run the principal minors of the
// twisted and the ordinary vertex matrix (fpm_wave_minors) while waves 2 and
// 3 load the diagonals.  Each FPM wave writes its minors into its own class's
// array a[class] as soon as it is done: its level lists sit in its own
// a[class] and dp[class] (order 10) or beyond the recursion's arrays (lower
// orders, whose lists outgrow 2^n doubles), never in the other class's a.
// After one barrier a[class][mask] becomes pm * pm (a[class][0] = 0) and
// pair_finish runs the recursion and assembly.  The minors never leave the
// workgroup: the same values, bit for bit, as k_fpm_wave + k_cdet_pair, with
// two launches and the minors' HBM round trip fewer.
// A configuration whose twisted or ordinary matrix trips the reference's FPM
// fallback (include/cdet/fast_principle_minors.h:180-186) is deferred: the
// minors of its other class go to pm, the flagged matrix to the fallback
// queue (fb_list / fb_count, as k_fpm_wave queues it) and the configuration
// to defer_list / defer_count; k_fallback_dets then writes the flagged
// matrix's pivoted determinants and k_cdet_pair's deferred form finishes the
// configuration.  (Pivoted determinants in registers here would need ~200
// VGPRs and cut the occupancy of every workgroup for the rare fallback.)
struct MinorsPairLds {
    // the FPM level lists fit in a[class] and dp[class]
    static __host__ __device__ constexpr bool lists_in_arrays(int n) {
        return fpm_wave_e0(n) <= (1 << n) && fpm_wave_e1(n) <= (1 << n);
    }
    static __host__ __device__ inline size_t arrays(int n, int mode) {
        return (mode == 2 ? 6 : 4) * (size_t(1) << n);
    }
    static __host__ __device__ inline size_t extra(int n) {
        return lists_in_arrays(n) ? 0 : 2 * static_cast<size_t>(fpm_wave_lds_doubles(n));
    }
    static __host__ __device__ inline size_t doubles(int n, int mode) {
        return arrays(n, mode) + 4 * static_cast<size_t>(n) + (size_t(1) << n) / 4 + 1 + extra(n);
    }
    static __host__ __device__ inline size_t bytes(int n, int mode) { return sizeof(double) * doubles(n, mode); }
};

// Four workgroups per CU (their LDS fills it) need at most 128 VGPRs each.
template <int N>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_minors_pair(
    KScalars ks, int n_rt, int mode, int items,
                                                     const int32_t* __restrict__ cfg_list,
                                                     const int64_t* __restrict__ cfg_pbase,
                                                     const int64_t* __restrict__ cfg_obase,
                                                     const double* __restrict__ g_up, const double* __restrict__ g_dn,
                                                     double* __restrict__ pm, double* __restrict__ out,
                                                     int32_t* __restrict__ status,
                                                     const uint16_t* __restrict__ masks_by_pop,
                                                     const int* __restrict__ pop_off, int32_t* __restrict__ fb_list,
                                                     int32_t* __restrict__ fb_count, int32_t* __restrict__ defer_list,
                                                     int32_t* __restrict__ defer_count) {
    constexpr int L = fpm_wave_L(N);
    constexpr int M = N - L;
    constexpr int T = 1 << L;
    constexpr uint32_t sc = 1u << N;
    constexpr int n2 = N * N;
    constexpr int E0 = fpm_wave_e0(N);
    constexpr bool kInArrays = MinorsPairLds::lists_in_arrays(N);
    extern __shared__ double lds_dyn[];
    __shared__ __attribute__((aligned(16))) double tb[kFoldWaves * kFoldSteps * 64];
    __shared__ int s_popoff[kMaxOrder + 2];
    __shared__ int s_bad[2];
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    double* a = lds_dyn;
    double* dp = a + 2 * sc;
    double* rhs = dp + 2 * sc;  // mode 2 only
    double* dU = lds_dyn + MinorsPairLds::arrays(N, mode);
    double* dD = dU + 2 * N;
    uint16_t* lmask = reinterpret_cast<uint16_t*>(dD + 2 * N);
    double* xtra = dD + 2 * N + sc / 4 + 1;
    // this wave's FPM level lists (waves 0 and 1)
    const int wc = wave & 1;
    double* list0 = kInArrays ? a + wc * sc : xtra + wc * fpm_wave_lds_doubles(N);
    double* list1 = kInArrays ? dp + wc * sc : list0 + E0;
    for (int k = tid; k < kMaxOrder + 2; k += nt) s_popoff[k] = pop_off[k];
    for (uint32_t m = tid; m < sc; m += nt) lmask[m] = masks_by_pop[m];
    for (int li = blockIdx.x; li < items; li += gridDim.x) {
        const int c = cfg_list[li];
        const int64_t pb = 2 * cfg_pbase[c];
        if (wave < 2) {
            double b[1 << M];
            // the order at run time (n_rt = N): with N folded in, the level loops
            // unroll completely and the kernel needs ~170 VGPRs instead of ~100
            const int r = fpm_wave_minors<M>(ks, n_rt, g_up + pb + static_cast<int64_t>(wave) * n2, ks.dup, list0,
                                             list1, lane, b);
            if (lane == 0) s_bad[wave] = r;
            if (r == kFpmWaveOk) {
                wave_sync_lds();  // the wave's list reads precede its writes of a[class]
                if (lane < T) {
                    double* ac = a + wave * sc;
#pragma unroll
                    for (int h = 0; h < (1 << M); ++h) ac[h * T + lane] = b[h];
                }
            }
        } else {
            pair_diagonals(ks, N, pb, g_up, g_dn, dU, dD, tid - 128, nt - 128);
        }
        __syncthreads();
        if ((s_bad[0] | s_bad[1]) != 0) {  // uniform
            if (wave < 2) {
                if (s_bad[wave] == kFpmWaveOk) {
                    double* dst = pm + static_cast<size_t>(4 * li + 2 * wave) * sc;
                    for (uint32_t m = lane; m < sc; m += 64) dst[m] = a[wave * sc + m];
                } else if (lane == 0) {
                    fb_list[atomicAdd(fb_count, 1)] = (li << 3) | (wave << 2) | 3;
                }
            }
            if (tid == 0) defer_list[atomicAdd(defer_count, 1)] = li;
            __syncthreads();  // the next configuration reuses LDS
            continue;
        }
        // a = pm_up * pm_down with identical spins, a[class][0] = 0
        for (uint32_t m = tid; m < 2 * sc; m += nt) {
            const double v = a[m];
            a[m] = (m & (sc - 1)) == 0 ? 0.0 : v * v;
        }
        __syncthreads();
        pair_finish(ks, n_rt, mode, dU, dD, a, dp, rhs, lmask, s_popoff, tb, out + cfg_obase[c],
                    status ? status + c : nullptr);
    }
}

// values[mask] -= 2 * ordinary_values[mask]
// (include/cdet/cdet_integrand_evaluator.h:348-350); with add set, the
// magnitude scales of the two classes add instead.
__global__ void k_combine_all(int64_t n, double* __restrict__ out, const double* __restrict__ out_ord, int add) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = add ? out[i] + 2.0 * out_ord[i] : out[i] - 2.0 * out_ord[i];
}

// twisted - DoubleDouble(2) * ordinary, rounded
// (include/cdet/cdet_integrand_evaluator.h:262, :559).
__global__ void k_combine_single(int64_t count, const int64_t* __restrict__ cfg_obase,
                                 const double* __restrict__ single_tmp, double* __restrict__ out) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= count) return;
    const DD tw{single_tmp[4 * c], single_tmp[4 * c + 1]};
    const DD od{single_tmp[4 * c + 2], single_tmp[4 * c + 3]};
    const DD r = dd_sub(tw, dd_mul(DD{2.0, 0.0}, od));
    out[cfg_obase[c]] = r.hi + r.lo;
}

// FP32 hi/lo split of a coefficient array: c = hi + lo + O(2^-48 |c|).
__global__ void k_split_f32(const double* __restrict__ c, int64_t n, float* __restrict__ hi, float* __restrict__ lo) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = c[i];
    const float h = static_cast<float>(x);
    hi[i] = h;
    lo[i] = static_cast<float>(x - static_cast<double>(h));
}

// Coefficients in the order the FP64 matrix-core contraction reads them.  A
// block (nu x nv, row-major) becomes (nu / 16) (nv / 16) steps of 256 values,
// step m (nv / 16) + t holding rows 16 m.., columns 16 t..; within a step,
// element (row i, column 4 r + g) sits at 64 r + 16 g + i, so the A operand
// of MFMA step r is one contiguous 512-byte read by lane 16 g + i.  One
// thread per output value (contiguous writes); the reads of a wave cover 16
// rows x 64 bytes.
__global__ void k_step_order(const double* __restrict__ c, int64_t n, int nu, int nv, double* __restrict__ out) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int64_t per = static_cast<int64_t>(nu) * nv;
    const int64_t b = e / per;
    const int w = static_cast<int>(e - b * per);
    const int step = w >> 8;
    const int k = w & 255;
    const int groups = nv >> 4;
    const int m = step / groups;
    const int t = step - m * groups;
    const int row = 16 * m + (k & 15);
    const int col = 16 * t + 4 * (k >> 6) + ((k >> 4) & 3);
    out[e] = c[b * per + static_cast<int64_t>(row) * nv + col];
}

// Fill kernel (graph-capture safe replacement of hipMemsetAsync).
// out[b] = sum over the nu*nv coefficients of block b of |c|: one wave per
// block (the table-value error scale of the error-bound kind).
__global__ __launch_bounds__(256) void k_block_abssum(const double* __restrict__ c, int64_t blocks, int per_block,
                                                      double* __restrict__ out) {
    const int64_t b = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (b >= blocks) return;  // uniform per wave
    const int lane = threadIdx.x & 63;
    const double* p = c + b * per_block;
    double s = 0.0;
    for (int i = lane; i < per_block; i += 64) s += fabs(p[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) out[b] = s;
}

// out[r] = sum over b of c[r nv + b] for every coefficient row r (block-major
// rows a of each block): one wave per row, lanes strided over b, then a fixed
// butterfly (row sums of the lower-branch tables, for lookups at v = 1).
__global__ __launch_bounds__(256) void k_row_sums(const double* __restrict__ c, int64_t rows, int nv,
                                                  double* __restrict__ out) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (r >= rows) return;  // uniform per wave
    const int lane = threadIdx.x & 63;
    const double* p = c + r * nv;
    double s = 0.0;
    for (int i = lane; i < nv; i += 64) s += p[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) out[r] = s;
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct cdet_batch_plan;

// Device buffer that only grows (see minors_scratch).
struct Scratch {
    void* ptr = nullptr;
    size_t cap = 0;
};

constexpr int kMinorLanes = 4;  // streams the order buckets of one plan run spread over
// Lanes of the plans of cdet_eval_batch_submit, whose two slots already run
// concurrently: two slots of two lanes fill the device's hardware queues (4
// per process by default).
constexpr int kSlotLanes = 2;

// Workspace of one plain (one configuration per launch) evaluation: its own
// stream and device buffers.  A context keeps a pool of them, so that host
// threads evaluating at the same time (the reference calls the integrand from
// one std::async thread per chain) each hold a workspace of their own.
struct PlainWork {
    hipStream_t stream = nullptr;
    double* d_verts = nullptr;
    double* d_gup = nullptr;
    double* d_gdn = nullptr;
    double* d_out = nullptr;
    double* d_out_ord = nullptr;
    double* d_single = nullptr;
    int32_t* d_small_i32 = nullptr;  // [0] status, [1] fallback queue length, [2..5] queue
    Scratch scratch;
};

// Scratch of cdet_green_pair_eval, grown to the largest call.  Each call's
// launches first wait (on the device) for the previous call's event, so calls
// on different streams never overlap on these buffers.
struct QueryScratch {
    int64_t cap = 0;
    int64_t tiles_cap = 0;
    int32_t nkeys = 0;
    hipEvent_t done = nullptr;  // recorded after each call's last launch
    double *gup = nullptr, *gdn = nullptr, *qu = nullptr, *qv = nullptr, *su = nullptr, *sv = nullptr;
    int32_t *qkey = nullptr, *qrank = nullptr, *sdest = nullptr, *hist = nullptr, *keyoff = nullptr;
    int32_t *tkey = nullptr, *tstart = nullptr, *tlen = nullptr, *ntiles = nullptr;
    int2* wgtot = nullptr;
};

struct cdet_ctx {
    cdet_params p{};
    int device = 0;
    int ncu = 256;  // compute units of the device
    uint64_t tables_gen = 0;  // bumped whenever the kernels' table pointers change
    hipStream_t stream = nullptr;
    DevTable tab[4];
    KScalars ks{};
    uint16_t* d_masks = nullptr;  // masks by popcount for every order, (kMaxOrder+1) * 2^kMaxOrder
    int* d_popoff = nullptr;      // (kMaxOrder+1) * (kMaxOrder+2)
    uint64_t* d_fallbacks = nullptr;
    unsigned long long* d_pair_cache = nullptr;  // Green-pair reuse: [0] reused, [1] looked up
    std::atomic<uint64_t> calls{0};  // counted from every calling thread
    int32_t* d_one = nullptr;   // cfg list {0}
    int64_t* d_zero64 = nullptr;  // pbase/obase {0}
    // plain-path workspaces: every one made (freed at destroy) and the idle ones
    std::mutex plain_mu;
    std::vector<PlainWork*> plain_all;
    std::vector<PlainWork*> plain_idle;
    // cdet_green_pair_eval
    std::mutex query_mu;
    QueryScratch query;
    // cdet_eval_batch_host (one call at a time per context)
    std::mutex host_mu;
    // derived table copies (step order, split FP32), made once per table
    std::mutex derive_mu;
    // launch-shape choices in effect (cdet_ctx_set_tuning)
    cdet_tuning tuning{};
    // Green-pair reuse of the driver's batch plans (cdet_ctx_set_pair_reuse)
    int pair_reuse = 1;
    // device address of this device's copy of `scratch` (no-allocation entry point)
    unsigned char* noalloc_scratch = nullptr;
    // host-buffer batch path (driver callback)
    int batch_mode = CDET_MODE_EXACT;
    int sum_order = CDET_SUM_TREE;  // cdet_ctx_set_sum_order
    cdet_batch_plan* host_plan = nullptr;
    std::vector<int32_t> host_plan_orders;
    // cdet_eval_batch_submit / wait: two slots, each with its own plan cache,
    // stream, device buffers, pinned host staging and launch thread (a launch
    // of a graph whose order buckets fork over several streams holds the
    // calling thread until the graph is about done, so the slot's copies and
    // graph launch run on the slot's own thread)
    struct HostSlot {
        std::thread worker;
        std::mutex mu;
        std::condition_variable cv;
        bool job = false;       // a launch is queued for the worker
        bool launched = true;   // the worker has finished the queued launch
        bool quit = false;
        int launch_rc = 0;
        std::string launch_err;
        cdet_batch_plan* plan = nullptr;
        std::vector<int32_t> orders;
        hipStream_t stream = nullptr;
        void* d_v = nullptr;
        size_t d_v_cap = 0;
        void* d_out = nullptr;
        size_t d_out_cap = 0;
        void* d_status = nullptr;
        size_t d_status_cap = 0;
        void* h_v = nullptr;  // pinned
        size_t h_v_cap = 0;
        void* h_out = nullptr;  // pinned
        size_t h_out_cap = 0;
        void* h_status = nullptr;  // pinned
        size_t h_status_cap = 0;
        int64_t count = 0;
        size_t out_bytes = 0;
        bool pending = false;
        hipStream_t lanes[kSlotLanes - 1] = {};  // fork streams of the slot's plans
    } slots[2];
    // fork streams of the other plans of the context, created on first use
    hipStream_t lanes[kMinorLanes - 1] = {};
    void* h_hv = nullptr;  // pinned staging of cdet_eval_batch_host
    size_t h_hv_cap = 0;
    void* h_hout = nullptr;
    size_t h_hout_cap = 0;
    std::map<const char*, size_t> host_bufs;  // cdet_host_alloc: pinned buffers by start, bytes (host_mu)
    void* h_hstatus = nullptr;
    size_t h_hstatus_cap = 0;
    void* d_hv = nullptr;
    size_t d_hv_cap = 0;
    void* d_hout = nullptr;
    size_t d_hout_cap = 0;
    void* d_hstatus = nullptr;
    size_t d_hstatus_cap = 0;
    // layout of the ordinary kinds now resident (CDET_ORDINARY_*)
    int ordinary_layout = CDET_ORDINARY_RESIDENT;
    // histogram-seed conditional probabilities (cdet_seed_cond_batch_host)
    void* d_seed_tab = nullptr;
    size_t seed_tab_cap = 0;
    uint64_t seed_hash = 0;
    void* d_seed_in = nullptr;
    size_t seed_in_cap = 0;
    void* d_seed_out = nullptr;
    size_t seed_out_cap = 0;
};

namespace {

void masks_by_popcount(std::vector<uint16_t>& masks, std::vector<int>& off) {
    const int M = 1 << kMaxOrder;
    masks.assign(static_cast<size_t>(kMaxOrder + 1) * M, 0);
    off.assign(static_cast<size_t>(kMaxOrder + 1) * (kMaxOrder + 2), 0);
    for (int n = 0; n <= kMaxOrder; ++n) {
        int w = 0;
        const int sc = 1 << n;
        for (int k = 0; k <= n + 1; ++k) {
            off[static_cast<size_t>(n) * (kMaxOrder + 2) + k] = w;
            if (k > n) break;
            for (int m = 0; m < sc; ++m)
                if (__builtin_popcount(static_cast<unsigned>(m)) == k) masks[static_cast<size_t>(n) * M + w++] = static_cast<uint16_t>(m);
        }
    }
}

// Double-double helpers on the host for the single-order prefactor
// (include/cdet/integrand_replica_randomly_sampled.h:156-160).
struct HDD {
    double hi, lo;
};
HDD hdd_mul(HDD x, HDD y) {
    const double p = x.hi * y.hi;
    double e = std::fma(x.hi, y.hi, -p);
    e += x.hi * y.lo + x.lo * y.hi + x.lo * y.lo;
    HDD r;
    r.hi = dd_quick_two_sum(p, e, r.lo);
    return r;
}
HDD hdd_add(HDD x, HDD y) {
    double s2, t2;
    double s1 = dd_two_sum(x.hi, y.hi, s2);
    const double t1 = dd_two_sum(x.lo, y.lo, t2);
    s2 += t1;
    s1 = dd_quick_two_sum(s1, s2, s2);
    s2 += t2;
    HDD r;
    r.hi = dd_quick_two_sum(s1, s2, r.lo);
    return r;
}
HDD hdd_sub(HDD x, HDD y) { return hdd_add(x, HDD{-y.hi, -y.lo}); }
HDD hdd_div(HDD x, HDD y) {
    const double q1 = x.hi / y.hi;
    const HDD r = hdd_sub(x, hdd_mul(HDD{q1, 0.0}, y));
    const double q2 = (r.hi + r.lo) / y.hi;
    HDD q = hdd_add(HDD{q1, 0.0}, HDD{q2, 0.0});
    const double s = q.hi + q.lo;
    q.lo = q.lo - (s - q.hi);
    q.hi = s;
    return q;
}

void refresh_scalars(cdet_ctx* ctx) {
    ++ctx->tables_gen;
    KScalars& k = ctx->ks;
    const cdet_params& p = ctx->p;
    k.N1 = p.N1;
    k.N2 = p.N2;
    k.nu = p.cheb_nu;
    k.nv = p.cheb_nv;
    k.SP = static_cast<int64_t>(4) * p.N1 * p.N2 * p.N1 * p.N2;
    k.beta = p.beta;
    k.U = p.U;
    k.branch_eps = p.branch_epsilon;
    k.tol = p.tol_fpm;
    k.shift = p.fpm_global_shift;
    const double eps0 = p.ztransform_epsilon0 > 0.0 ? p.ztransform_epsilon0 : 2.0e-16;
    k.thr = eps0 * p.ztransform_safety;
    k.filter = k.thr > 0.0 ? 1 : 0;
    k.scale = p.scale;
    k.dup = p.diag_up_shift;
    k.ddn = p.diag_down_shift;
    k.counter = p.diag_up_shift * p.diag_down_shift;
    for (int cls = 0; cls < 2; ++cls) {
        const int rc = cls == 0 ? 2 : 1;
        k.c1[cls] = static_cast<double>(-rc) * p.U * p.beta;
        for (int order = 0; order <= kMaxOrder; ++order) {
            long double num = 1.0L;
            for (int i = 0; i < order; ++i) num *= -static_cast<long double>(rc);
            long double den = static_cast<long double>(p.N1) * static_cast<long double>(p.N2);
            for (int i = 2; i <= order; ++i) den *= static_cast<long double>(i);
            k.pref[cls][order] = static_cast<double>(num / den);
            double dnum = 1.0;
            for (int i = 0; i < order; ++i) dnum *= -static_cast<double>(rc);
            HDD fact{1.0, 0.0};
            for (int i = 1; i <= order; ++i) fact = hdd_mul(fact, HDD{static_cast<double>(i), 0.0});
            const HDD dd = hdd_mul(hdd_mul(fact, HDD{static_cast<double>(p.N1), 0.0}), HDD{static_cast<double>(p.N2), 0.0});
            const HDD q = hdd_div(HDD{dnum, 0.0}, dd);
            k.pref_hi[cls][order] = q.hi;
            k.pref_lo[cls][order] = q.lo;
        }
    }
    for (int kind = 0; kind < 4; ++kind) {
        const DevTable& t = ctx->tab[kind].alias >= 0 ? ctx->tab[ctx->tab[kind].alias] : ctx->tab[kind];
        k.coeff[kind][0] = t.upper;
        k.coeff[kind][1] = t.lower;
        k.corner[kind][0] = t.ucorner;
        k.corner[kind][1] = t.lcorner;
        k.steps[kind][0] = t.steps[0];
        k.steps[kind][1] = t.steps[1];
        k.split_hi[kind][0] = t.split[0][0];
        k.split_lo[kind][0] = t.split[0][1];
        k.split_hi[kind][1] = t.split[1][0];
        k.split_lo[kind][1] = t.split[1][1];
        k.abssum[kind][0] = t.abssum[0];
        k.abssum[kind][1] = t.abssum[1];
        k.rowsum[kind] = t.rowsum;
    }
    for (int order = 0; order <= kMaxOrder; ++order) {
        k.group16_R[order] = ctx->tuning.group16_min_terms[order];
        k.group64_R[order] = ctx->tuning.group64_min_terms[order];
        k.tree_bits[order] = 0;
        for (int layer = 0; layer <= kMaxOrder; ++layer)
            k.tree_bits[order] |= static_cast<uint64_t>(tree_lanes_log2(order, layer)) << (3 * layer);
    }
    k.tree = 0;  // the reference's order; a launch in a non-exact mode sets it on its copy
    for (int cls = 0; cls < 2; ++cls) {
        const int ku = cls == 0 ? 0 : 2;
        k.same_spin[cls] = (k.coeff[ku][0] == k.coeff[ku + 1][0] && k.coeff[ku][1] == k.coeff[ku + 1][1] &&
                            k.corner[ku][0] == k.corner[ku + 1][0] && k.corner[ku][1] == k.corner[ku + 1][1] &&
                            k.dup == k.ddn)
                               ? 1
                               : 0;
    }
}

bool tables_ready(const cdet_ctx* ctx) {
    for (int kind = 0; kind < 4; ++kind)
        if (!ctx->ks.coeff[kind][0] || !ctx->ks.coeff[kind][1] || !ctx->ks.corner[kind][0] || !ctx->ks.corner[kind][1])
            return false;
    return true;
}

int grow(void** p, size_t& cap, size_t need) {
    if (need <= cap) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    HIP_TRY(hipMalloc(p, need));
    cap = need;
    return 0;
}

void plain_work_free(PlainWork* w) {
    if (!w) return;
    if (w->stream) {
        (void)hipStreamSynchronize(w->stream);
        (void)hipStreamDestroy(w->stream);
    }
    void* bufs[] = {w->d_verts, w->d_gup, w->d_gdn, w->d_out, w->d_out_ord, w->d_single, w->d_small_i32,
                    w->scratch.ptr};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    delete w;
}

// A plain-path workspace sized for CDET_MAX_ORDER (the minors scratch grows
// on first use of an order).
int plain_work_make(PlainWork** out) {
    auto* w = new PlainWork();
    const int Mx = kMaxOrder;
    auto made = [&]() -> int {
        HIP_TRY(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
        HIP_TRY(hipMalloc(&w->d_verts, sizeof(double) * 4 * Mx));
        HIP_TRY(hipMalloc(&w->d_gup, sizeof(double) * 2 * Mx * Mx));
        HIP_TRY(hipMalloc(&w->d_gdn, sizeof(double) * 2 * Mx * Mx));
        HIP_TRY(hipMalloc(&w->d_out, sizeof(double) << Mx));
        HIP_TRY(hipMalloc(&w->d_out_ord, sizeof(double) << Mx));
        HIP_TRY(hipMalloc(&w->d_single, sizeof(double) * 4));
        HIP_TRY(hipMalloc(&w->d_small_i32, sizeof(int32_t) * 8));
        return 0;
    };
    const int rc = made();
    if (rc) {
        plain_work_free(w);
        return rc;
    }
    *out = w;
    return 0;
}

// Take an idle workspace of the context, or make one.
int plain_work_acquire(cdet_ctx* ctx, PlainWork** out) {
    {
        std::lock_guard<std::mutex> lk(ctx->plain_mu);
        if (!ctx->plain_idle.empty()) {
            *out = ctx->plain_idle.back();
            ctx->plain_idle.pop_back();
            return 0;
        }
    }
    PlainWork* w = nullptr;
    const int rc = plain_work_make(&w);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(ctx->plain_mu);
    ctx->plain_all.push_back(w);
    *out = w;
    return 0;
}

void plain_work_release(cdet_ctx* ctx, PlainWork* w) {
    std::lock_guard<std::mutex> lk(ctx->plain_mu);
    ctx->plain_idle.push_back(w);
}

// Free the buffers once the last call that used them has finished.
void query_scratch_free(QueryScratch& q) {
    if (q.done) (void)hipEventSynchronize(q.done);
    void* bufs[] = {q.gup, q.gdn, q.qu, q.qv, q.su, q.sv, q.qkey, q.qrank, q.sdest, q.hist, q.keyoff,
                    q.tkey, q.tstart, q.tlen, q.ntiles, q.wgtot};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    hipEvent_t done = q.done;
    q = QueryScratch();
    q.done = done;
}

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of the no-allocation entry point (cdet_eval_all_subsets_noalloc):
// every buffer one configuration of order <= CDET_MAX_ORDER needs beyond its
// vertices and values.  Each device holds its own copy for the life of the
// process.  Host code never uses the name: the device address of the current
// device's copy is looked up once per device (hipGetSymbolAddress) when the
// first context is created on it and kept per device.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocBytes = size_t(4) << 20;  // 4 MiB; order 14 needs 3.1 MB
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
unsigned char* g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

// Address of `scratch` on `device`, which must be the current device.
static int noalloc_scratch_of(int device, unsigned char** out) {
    if (device < 0 || device >= kMaxDevices) return fail(CDET_ERR_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(CDET_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_noalloc_ptr[device] = static_cast<unsigned char*>(ptr);
    }
    *out = g_noalloc_ptr[device];
    return 0;
}

static constexpr size_t kMaxLds = 160 * 1024 - 2048;
// Workgroups of k_fallback_dets (single waves that loop over the queue's
// work items; its launches are mostly empty, fallbacks being rare).
static constexpr int kDetGridRows = 256;
static constexpr int kDeferGrid = 16;  // workgroups of the deferred k_cdet_pair pass
static constexpr int kGscrGrid = 1024;  // workgroups of the global-scratch forms
static bool fpm_gscr(int n) { return FpmLds::bytes(n) > kMaxLds; }
// k_cdet's static LDS (wave-wide chain terms) comes out of its dynamic budget.
static constexpr size_t kCdetMaxDyn = kMaxLds - sizeof(double) * kFoldWaves * kFoldSteps * 64;
// Orders from which k_fpm and k_cdet run 1024-thread workgroups: their LDS
// workspaces leave room for one workgroup per CU (or they work in global
// scratch), and 256 threads would leave each SIMD a single wave.  At order 12
// (two workgroups per CU) the wider form measured slower.
static constexpr int kWideOrder = 13;

// Built-in launch shapes (cdet_tuning defaults).
static cdet_tuning default_tuning() {
    cdet_tuning t{};
    t.contract_wg_per_cu = 8;
    t.contract_tile = 0;
    t.pair_min_per_cu = 2;
    t.max_lanes = 4;
    for (int n = 0; n <= kMaxOrder; ++n) {
        t.minors_threads[n] = n >= kWideOrder ? 1024 : 256;
        t.group16_min_terms[n] = 64;
        t.group64_min_terms[n] = 128;
    }
    t.fused_minors_max_order = 1;
    return t;
}
static bool wide_minors(const cdet_ctx* ctx, int n) { return ctx->tuning.minors_threads[n] >= 1024; }
static bool cdet_gscr(int n, int mode) { return CdetLds::bytes(n, mode) > kCdetMaxDyn; }

// Outputs and queues of one minors launch.
struct MinorsOut {
    double* out;
    double* out_ord;
    double* single_tmp;
    int32_t* status;
    int32_t* fb_list;   // capacity 4 * count entries
    int32_t* fb_count;
    bool paired = false;  // modes 0, 2: k_cdet takes both classes of a configuration
    const double* err_up = nullptr;  // mode 3: table-value error scale per Green pair
    const double* err_dn = nullptr;
    int32_t* defer_list = nullptr;   // paired plans: k_minors_pair's deferred configurations (capacity count)
    int32_t* defer_count = nullptr;  // two words, zero between runs
};

// Global scratch of one order bucket: the principal minors of every spin
// matrix (4 per configuration), then the k_fpm and k_cdet workspaces of the
// global-scratch forms.  The buffer is owned by the caller (a plan, or the
// context for the plain form) and only grows; a plan sizes it at creation so
// that its captured graph never sees a reallocation.
static size_t minors_scratch_bytes(int n, int count, int mode) {
    if (mode == 3)  // error-bound kind: one workspace per workgroup of k_bound
        return sizeof(double) * bound_doubles(n) * std::min<size_t>(2 * static_cast<size_t>(count), kBoundGrid);
    const size_t sc = size_t(1) << n;
    size_t b = sizeof(double) * 4 * static_cast<size_t>(count) * sc;
    if (fpm_gscr(n)) b += FpmLds::bytes(n) * std::min<size_t>(4 * static_cast<size_t>(count), kGscrGrid);
    if (cdet_gscr(n, mode)) b += CdetLds::bytes(n, mode) * std::min<size_t>(2 * static_cast<size_t>(count), kGscrGrid);
    return b;
}

// Runs of several orders: orders 2 .. kMixedOrderMax, whose principal minors
// all take the one-wave kernel (k_fpm_wave), and whose fallback determinants
// all run in registers.  Up to order kMixedSmallMax, n - fpm_wave_L(n) = 1, so
// the run's minors need no per-configuration subtree depth.
constexpr int kMixedOrderMax = 10;
constexpr int kMixedSmallMax = 7;
static_assert(kMixedOrderMax <= kWaveFpmMaxOrder && kMixedOrderMax <= kDetRegMax && kMixedOrderMax <= kPairOrderMax,
              "a mixed run takes the one-wave minors, register determinants and paired recursion");

// Launch the minors and connected-value kernels of one order bucket, or, with
// cfg_order non-null, of a run of configurations of orders 2 .. n (n <=
// kMixedOrderMax, modes 0 - 2): configuration c has order cfg_order[c], and
// the kernels lay out every configuration's minors at the stride of order n.
static int launch_minors(cdet_ctx* ctx, Scratch& scratch, int n, int mode, const int32_t* cfg_list, int count,
                         const int64_t* pbase, const int64_t* obase, const double* gup, const double* gdn,
                         const MinorsOut& o, hipStream_t stream, const int32_t* cfg_order = nullptr,
                         bool tree = false) {
    if (count <= 0) return 0;
    // tree: the recursion's sums in tree order (tree_layer), for the modes
    // whose values are not bit-identical to the reference
    KScalars ks = ctx->ks;
    ks.tree = tree ? 1 : 0;
    if (cfg_order && (n < 2 || n > kMixedOrderMax || mode == 3))
        return fail(CDET_ERR_ARGUMENT, "launch_minors: a run of mixed orders needs orders 2..7 and modes 0..2");
    int rc = grow(&scratch.ptr, scratch.cap, minors_scratch_bytes(n, count, mode));
    if (rc) return rc;
    const size_t sc = size_t(1) << n;
    if (mode == 3) {  // error bounds of the all-subset values
        const uint16_t* masks3 = ctx->d_masks + static_cast<size_t>(n) * (1u << kMaxOrder);
        const int* popoff3 = ctx->d_popoff + n * (kMaxOrder + 2);
        const int items = 2 * count;
        const int grid = std::min(items, kBoundGrid);
        hipLaunchKernelGGL(k_bound, dim3(grid), dim3(kBoundThreads), 0, stream, ks, n, items, cfg_list, pbase,
                           obase, gup, gdn, o.err_up, o.err_dn, o.out, o.out_ord, masks3, popoff3,
                           static_cast<double*>(scratch.ptr), bound_doubles(n));
        HIP_TRY(hipGetLastError());
        return 0;
    }
    double* pm = static_cast<double*>(scratch.ptr);
    double* next = pm + 4 * static_cast<size_t>(count) * sc;
    const uint16_t* masks = ctx->d_masks + static_cast<size_t>(n) * (1u << kMaxOrder);
    const int* popoff = ctx->d_popoff + n * (kMaxOrder + 2);
    const bool fused = !cfg_order && o.paired && o.defer_count && (mode == 0 || mode == 2) && n >= 2 &&
                       n <= std::min(kPairOrderMax, ctx->tuning.fused_minors_max_order) && ctx->ks.same_spin[0] &&
                       ctx->ks.same_spin[1];
    if (fused) {
        // minors and recursion in one workgroup per configuration; the
        // configurations whose minors need the pivoted fallback finish in
        // k_fallback_dets and the deferred k_cdet_pair
        const size_t lds = MinorsPairLds::bytes(n, mode);
#define CDET_MINORS_PAIR(NN)                                                                                          \
    case NN:                                                                                                          \
        hipLaunchKernelGGL(k_minors_pair<NN>, dim3(count), dim3(256), lds, stream, ks, NN, mode, count,         \
                           cfg_list,                                                                                  \
                           pbase, obase, gup, gdn, pm, o.out, o.status, masks, popoff, o.fb_list, o.fb_count,         \
                           o.defer_list, o.defer_count);                                                              \
        break;
        switch (n) {
            CDET_MINORS_PAIR(2)
            CDET_MINORS_PAIR(3)
            CDET_MINORS_PAIR(4)
            CDET_MINORS_PAIR(5)
            CDET_MINORS_PAIR(6)
            CDET_MINORS_PAIR(7)
            CDET_MINORS_PAIR(8)
            CDET_MINORS_PAIR(9)
            default:
                CDET_MINORS_PAIR(10)
        }
#undef CDET_MINORS_PAIR
        const size_t det_lds = sizeof(double) * static_cast<size_t>(n) * n;
        hipLaunchKernelGGL(k_fallback_dets<kDetRegMax>, dim3(kDetGridRows), dim3(kDetThreads), det_lds, stream, ks, n,
                           cfg_list, pbase, gup, gdn, o.fb_list, o.fb_count, pm, o.status, ctx->d_fallbacks, nullptr,
                           n);
        hipLaunchKernelGGL(k_cdet_pair, dim3(kDeferGrid), dim3(256), CdetPairLds::bytes(n, mode), stream, ks, n,
                           mode, 0, cfg_list, pbase, obase, gup, gdn, pm, o.out, o.status, ctx->d_masks, ctx->d_popoff,
                           o.fb_count, o.defer_list, o.defer_count, nullptr, n);
        HIP_TRY(hipGetLastError());
        return 0;
    }
    if (n >= 2) {
        const int items = (ctx->ks.same_spin[0] && ctx->ks.same_spin[1] ? 2 : 4) * count;
        if (cfg_order && n > kMixedSmallMax) {
            // subtree depth per configuration
            const size_t lds = sizeof(double) * static_cast<size_t>(fpm_wave_lds_doubles(n));
            hipLaunchKernelGGL(k_fpm_wave<0>, dim3(items), dim3(64), lds, stream, ks, n, items, cfg_list, pbase,
                               gup, gdn, pm, o.fb_list, o.fb_count, cfg_order, n);
        } else if (n <= kWaveFpmMaxOrder) {
            // one wavefront per spin matrix, no workgroup barriers
            const size_t lds = sizeof(double) * static_cast<size_t>(fpm_wave_lds_doubles(n));
            switch (n - fpm_wave_L(n)) {
                case 1:
                    hipLaunchKernelGGL(k_fpm_wave<1>, dim3(items), dim3(64), lds, stream, ks, n, items, cfg_list,
                                       pbase, gup, gdn, pm, o.fb_list, o.fb_count, cfg_order, n);
                    break;
                case 2:
                    hipLaunchKernelGGL(k_fpm_wave<2>, dim3(items), dim3(64), lds, stream, ks, n, items, cfg_list,
                                       pbase, gup, gdn, pm, o.fb_list, o.fb_count, cfg_order, n);
                    break;
                case 3:
                    hipLaunchKernelGGL(k_fpm_wave<3>, dim3(items), dim3(64), lds, stream, ks, n, items, cfg_list,
                                       pbase, gup, gdn, pm, o.fb_list, o.fb_count, cfg_order, n);
                    break;
                default:
                    hipLaunchKernelGGL(k_fpm_wave<4>, dim3(items), dim3(64), lds, stream, ks, n, items, cfg_list,
                                       pbase, gup, gdn, pm, o.fb_list, o.fb_count, cfg_order, n);
                    break;
            }
        } else if (fpm_gscr(n)) {
            const int grid = std::min(items, kGscrGrid);
            if (wide_minors(ctx, n))
                hipLaunchKernelGGL((k_fpm<true, 1024>), dim3(grid), dim3(1024), 0, stream, ks, n, items, cfg_list,
                                   pbase, gup, gdn, pm, o.fb_list, o.fb_count, next, FpmLds::doubles(n));
            else
                hipLaunchKernelGGL((k_fpm<true, 256>), dim3(grid), dim3(256), 0, stream, ks, n, items, cfg_list,
                                   pbase, gup, gdn, pm, o.fb_list, o.fb_count, next, FpmLds::doubles(n));
            next += static_cast<size_t>(grid) * FpmLds::doubles(n);
        } else {
            if (wide_minors(ctx, n))
                hipLaunchKernelGGL((k_fpm<false, 1024>), dim3(items), dim3(1024), FpmLds::bytes(n), stream, ks, n,
                                   items, cfg_list, pbase, gup, gdn, pm, o.fb_list, o.fb_count, nullptr, 0);
            else
                hipLaunchKernelGGL((k_fpm<false, 256>), dim3(items), dim3(256), FpmLds::bytes(n), stream, ks, n,
                                   items, cfg_list, pbase, gup, gdn, pm, o.fb_list, o.fb_count, nullptr, 0);
        }
        const size_t det_lds = sizeof(double) * static_cast<size_t>(n) * n;
        if (n <= kDetRegMax)
            hipLaunchKernelGGL(k_fallback_dets<kDetRegMax>, dim3(kDetGridRows), dim3(kDetThreads), det_lds, stream, ks,
                               n, cfg_list, pbase, gup, gdn, o.fb_list, o.fb_count, pm, o.status, ctx->d_fallbacks,
                               cfg_order, n);
        else
            hipLaunchKernelGGL(k_fallback_dets<kMaxOrder>, dim3(kDetGridRows), dim3(kDetThreads), det_lds, stream, ks,
                               n, cfg_list, pbase, gup, gdn, o.fb_list, o.fb_count, pm, o.status, ctx->d_fallbacks,
                               cfg_order, n);
    }
    const int paired = (mode != 1 && o.paired) ? 1 : 0;
    if (paired && n <= kPairOrderMax) {
        hipLaunchKernelGGL(k_cdet_pair, dim3(count), dim3(256), CdetPairLds::bytes(n, mode), stream, ks, n, mode,
                           count, cfg_list, pbase, obase, gup, gdn, pm, o.out, o.status, ctx->d_masks, ctx->d_popoff,
                           o.fb_count, nullptr, nullptr, cfg_order, n);
        HIP_TRY(hipGetLastError());
        return 0;
    }
    const int items = paired ? count : 2 * count;
    if (cdet_gscr(n, mode)) {
        const int grid = std::min(items, kGscrGrid);
        // connected values in LDS, the products in global scratch, where they fit
        const bool split = mode != 2 && CdetSplitLds::bytes(n) <= kCdetMaxDyn;
        const size_t lds = split ? CdetSplitLds::bytes(n) : 0;
#define CDET_GSCR_LAUNCH(LAY, NT)                                                                                    \
    hipLaunchKernelGGL((k_cdet<LAY, NT>), dim3(grid), dim3(NT), lds, stream, ks, n, mode, items, cfg_list, pbase,     \
                       obase, gup, gdn, pm, o.out, o.out_ord, o.single_tmp, o.status, ctx->d_masks, ctx->d_popoff,   \
                       next, CdetLds::doubles(n, mode), o.fb_count, paired, cfg_order, n)
        if (wide_minors(ctx, n)) {
            if (split)
                CDET_GSCR_LAUNCH(kCdetLaySplit, 1024);
            else
                CDET_GSCR_LAUNCH(kCdetLayGlobal, 1024);
        } else {
            if (split)
                CDET_GSCR_LAUNCH(kCdetLaySplit, 256);
            else
                CDET_GSCR_LAUNCH(kCdetLayGlobal, 256);
        }
#undef CDET_GSCR_LAUNCH
    } else {
        if (wide_minors(ctx, n))
            hipLaunchKernelGGL((k_cdet<kCdetLayLds, 1024>), dim3(items), dim3(1024), CdetLds::bytes(n, mode), stream,
                               ks, n, mode, items, cfg_list, pbase, obase, gup, gdn, pm, o.out, o.out_ord,
                               o.single_tmp, o.status, ctx->d_masks, ctx->d_popoff, nullptr, 0, o.fb_count, paired,
                               cfg_order, n);
        else
            hipLaunchKernelGGL((k_cdet<kCdetLayLds, 256>), dim3(items), dim3(256), CdetLds::bytes(n, mode), stream,
                               ks, n, mode, items, cfg_list, pbase, obase, gup, gdn, pm, o.out, o.out_ord,
                               o.single_tmp, o.status, ctx->d_masks, ctx->d_popoff, nullptr, 0, o.fb_count, paired,
                               cfg_order, n);
    }

    HIP_TRY(hipGetLastError());
    return 0;
}

static int launch_combine(int mode, int64_t count, int64_t out_total, const int64_t* obase, const MinorsOut& o,
                          hipStream_t stream) {
    if (mode == 1) {
        hipLaunchKernelGGL(k_combine_single, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0, stream,
                           count, obase, o.single_tmp, o.out);
    } else if (!o.paired) {  // paired runs combined the classes in k_cdet
        hipLaunchKernelGGL(k_combine_all, dim3(static_cast<unsigned>((out_total + 255) / 256)), dim3(256), 0, stream,
                           out_total, o.out, o.out_ord, mode >= 2 ? 1 : 0);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// The 64-query FP64 kernel needs both coefficient counts to be multiples of
// 16; other shapes use the 32-query kernel.
static bool use_f64_v2(const KScalars& ks) { return ks.nu % 16 == 0 && ks.nv % 16 == 0; }
// Lookups per tile.  The FP64 matrix-core contraction takes tiles of up to
// kF64Tile lookups when a batch has enough of them to give every resident
// wave (4 per SIMD) about three tiles; smaller batches take tiles of 16
// lookups, so that more, smaller tiles spread the matrix-core work evenly
// over the SIMDs (each extra tile re-reads its coefficient block, which the
// large batches cannot afford).
static int mfma_tile_queries(const KScalars& ks, int mode, int64_t lookups, int ncu, int forced = 0) {
    if (mode != CDET_MODE_MFMA64 || !use_f64_v2(ks)) return kMfmaTile;
    if (forced == 16 || forced == kF64Tile) return forced;
    const int64_t resident_waves = 16 * static_cast<int64_t>(ncu);
    return lookups / kF64Tile >= 3 * resident_waves ? kF64Tile : 16;
}

static int table_sizes(const cdet_ctx* ctx, size_t& ncoef, size_t& ncorner);

// Step-order copies of every resident table for the FP64 matrix-core
// contraction (nu, nv multiples of 16), built once per table.
static int ensure_step_tables(cdet_ctx* ctx) {
    if (!use_f64_v2(ctx->ks)) return 0;
    std::lock_guard<std::mutex> lock(ctx->derive_mu);
    size_t nc, nk;
    table_sizes(ctx, nc, nk);
    bool made = false;
    for (int kind = 0; kind < 4; ++kind) {
        DevTable& t = ctx->tab[kind];
        if (t.alias >= 0 || !t.upper) continue;
        for (int br = 0; br < 2; ++br) {
            if (t.steps[br]) continue;
            HIP_TRY(hipMalloc(&t.steps[br], nc * sizeof(double)));
            hipLaunchKernelGGL(k_step_order, dim3(static_cast<unsigned>((nc + 255) / 256)), dim3(256), 0, ctx->stream,
                               br == 0 ? t.upper : t.lower, static_cast<int64_t>(nc), ctx->p.cheb_nu, ctx->p.cheb_nv,
                               t.steps[br]);
            HIP_TRY(hipGetLastError());
            made = true;
        }
    }
    if (made) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        refresh_scalars(ctx);
    }
    return 0;
}

// FP32 hi/lo copies of every resident table (split-FP32 mode), built once.
static int ensure_split_tables(cdet_ctx* ctx) {
    std::lock_guard<std::mutex> lock(ctx->derive_mu);
    size_t nc, nk;
    const size_t S = static_cast<size_t>(ctx->p.N1) * ctx->p.N2;
    nk = 4 * S * S;
    nc = nk * static_cast<size_t>(ctx->p.cheb_nu) * ctx->p.cheb_nv;
    bool made = false;
    for (int kind = 0; kind < 4; ++kind) {
        DevTable& t = ctx->tab[kind];
        if (t.alias >= 0 || !t.upper) continue;
        for (int br = 0; br < 2; ++br) {
            if (t.split[br][0]) continue;
            HIP_TRY(hipMalloc(&t.split[br][0], nc * sizeof(float)));
            HIP_TRY(hipMalloc(&t.split[br][1], nc * sizeof(float)));
            const double* src = br == 0 ? t.upper : t.lower;
            hipLaunchKernelGGL(k_split_f32, dim3(static_cast<unsigned>((nc + 255) / 256)), dim3(256), 0, ctx->stream, src,
                               static_cast<int64_t>(nc), t.split[br][0], t.split[br][1]);
            HIP_TRY(hipGetLastError());
            made = true;
        }
    }
    if (made) {  // only a change of the table pointers re-captures the plans' graphs
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        refresh_scalars(ctx);
    }
    return 0;
}

// Per-block coefficient magnitude sums of every resident table (the
// error-bound kind), built once per table.
static int ensure_abssum_tables(cdet_ctx* ctx) {
    std::lock_guard<std::mutex> lock(ctx->derive_mu);
    size_t nc, nk;
    table_sizes(ctx, nc, nk);
    bool made = false;
    for (int kind = 0; kind < 4; ++kind) {
        DevTable& t = ctx->tab[kind];
        if (t.alias >= 0 || !t.upper) continue;
        for (int br = 0; br < 2; ++br) {
            if (t.abssum[br]) continue;
            HIP_TRY(hipMalloc(&t.abssum[br], nk * sizeof(double)));
            hipLaunchKernelGGL(k_block_abssum, dim3(static_cast<unsigned>((nk + 3) / 4)), dim3(256), 0, ctx->stream,
                               br == 0 ? t.upper : t.lower, static_cast<int64_t>(nk),
                               ctx->p.cheb_nu * ctx->p.cheb_nv, t.abssum[br]);
            HIP_TRY(hipGetLastError());
            made = true;
        }
    }
    if (made) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        refresh_scalars(ctx);
    }
    return 0;
}

// Row sums of the lower-branch coefficient blocks of every resident table,
// built once per table for the FP64 matrix-core mode: a lookup at v = 1
// (every diagonal pair, i = j, and any pair of equal times) has T_b(v) = 1 for
// all b, so its value is sum_a T_a(u) sum_b c_ab, a Clenshaw over one row of
// sums instead of a block contraction.
static int ensure_rowsum_tables(cdet_ctx* ctx) {
    std::lock_guard<std::mutex> lock(ctx->derive_mu);
    size_t nc, nk;
    table_sizes(ctx, nc, nk);
    const int64_t rows = static_cast<int64_t>(nc / static_cast<size_t>(ctx->p.cheb_nv));
    bool made = false;
    for (int kind = 0; kind < 4; ++kind) {
        DevTable& t = ctx->tab[kind];
        if (t.alias >= 0 || !t.lower || t.rowsum) continue;
        HIP_TRY(hipMalloc(&t.rowsum, static_cast<size_t>(rows) * sizeof(double)));
        hipLaunchKernelGGL(k_row_sums, dim3(static_cast<unsigned>((rows + 3) / 4)), dim3(256), 0, ctx->stream, t.lower,
                           rows, ctx->p.cheb_nv, t.rowsum);
        HIP_TRY(hipGetLastError());
        made = true;
    }
    if (made) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        refresh_scalars(ctx);
    }
    return 0;
}

static int contract_launch_mfma(const KScalars& ks, int mode, int grid, const int32_t* ntiles, const int32_t* tkey,
                                const int32_t* tstart, const int32_t* tlen, const double* su, const double* sv,
                                const int32_t* sdest, double* gup, double* gdn, hipStream_t s, int tile) {
    if (mode == CDET_MODE_MFMA64 && use_f64_v2(ks)) {
        const bool same = ks.same_spin[0] && ks.same_spin[1];
        const int chunk = tile <= 16 ? contract_f64_slab_chunk(ks.nu >> 4, same) : 0;
        if (chunk && same)
            hipLaunchKernelGGL((k_contract_f64_slabs<true, 2>), dim3(grid), dim3(256), 0, s, ks, ntiles, tkey, tstart,
                               tlen, su, sv, sdest, gup, gdn);
        else if (chunk)
            hipLaunchKernelGGL((k_contract_f64_slabs<false, 2>), dim3(grid), dim3(256), 0, s, ks, ntiles, tkey, tstart,
                               tlen, su, sv, sdest, gup, gdn);
        else if (tile <= 16 && same)
            hipLaunchKernelGGL((k_contract_f64<true, 1>), dim3(grid), dim3(256), 0, s, ks, ntiles, tkey, tstart, tlen,
                               su, sv, sdest, gup, gdn);
        else if (tile <= 16)
            hipLaunchKernelGGL((k_contract_f64<false, 1>), dim3(grid), dim3(256), 0, s, ks, ntiles, tkey, tstart,
                               tlen, su, sv, sdest, gup, gdn);
        else if (same)
            hipLaunchKernelGGL((k_contract_f64<true, 4>), dim3(grid), dim3(256), 0, s, ks, ntiles, tkey, tstart, tlen,
                               su, sv, sdest, gup, gdn);
        else
            hipLaunchKernelGGL((k_contract_f64<false, 4>), dim3(grid), dim3(256), 0, s, ks, ntiles, tkey, tstart,
                               tlen, su, sv, sdest, gup, gdn);
    } else if (mode == CDET_MODE_MFMA64) {
        hipLaunchKernelGGL(k_contract_mfma<CDET_MODE_MFMA64>, dim3(grid), dim3(256), 0, s, ks, ntiles, tkey, tstart,
                           tlen, su, sv, sdest, gup, gdn);
    } else {
        if (!ks.split_hi[0][0] || !ks.split_hi[2][0]) return fail(CDET_ERR_TABLES, "split-FP32 tables were not built");
        hipLaunchKernelGGL(k_contract_mfma<CDET_MODE_SPLIT32>, dim3(grid), dim3(256), 0, s, ks, ntiles, tkey, tstart,
                           tlen, su, sv, sdest, gup, gdn);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// Grid of a persistent contraction launch: one wave per tile up to 8
// workgroups per CU, a multiple of the 8 XCDs.
static int contract_grid(const cdet_ctx* ctx, int64_t max_tiles, int mode = CDET_MODE_EXACT, int tile = 64) {
    int64_t per_cu = ctx->tuning.contract_wg_per_cu;
    if (mode == CDET_MODE_MFMA64 && use_f64_v2(ctx->ks)) {
        // no more workgroups than fit at once (256 threads = one wave per SIMD)
        const bool same = ctx->ks.same_spin[0] && ctx->ks.same_spin[1];
        const int chunk = tile <= 16 ? contract_f64_slab_chunk(ctx->ks.nu >> 4, same) : 0;
        int fit = tile <= 16 ? (same ? contract_f64_waves<true, 1>() : contract_f64_waves<false, 1>())
                             : (same ? contract_f64_waves<true, 4>() : contract_f64_waves<false, 4>());
        if (chunk) fit = same ? contract_f64_slab_waves<true, 2>() : contract_f64_slab_waves<false, 2>();
        per_cu = std::min<int64_t>(per_cu, fit);
    }
    int64_t g = std::min<int64_t>(std::max<int64_t>((max_tiles + 3) / 4, 1), static_cast<int64_t>(ctx->ncu) * per_cu);
    g = ((g + 7) / 8) * 8;
    return static_cast<int>(g);
}

static int alloc_table(cdet_ctx* ctx, int kind);
static void invalidate_split(DevTable& t);

// A stream on a hardware queue of its own, over every CU of the device.  HIP
// multiplexes ordinary streams over a small pool of hardware queues
// (GPU_MAX_HW_QUEUES, 4 by default), so two streams of a process can share a
// queue and then run their kernels one after the other; a stream made with a
// CU mask is given a queue of its own.  Batches in flight overlap only on such
// streams (order-10 batches of 3360 configurations, 4 in flight: 7.2 M
// configurations/s against 5.5 M on ordinary streams).  The stream
// synchronizes with the null stream like a stream made by hipStreamCreate.
// Only the caller's in-flight streams are made this way: the context stream,
// the host slots and the order lanes stay ordinary non-blocking streams,
// which measured faster for the host-slot sweeps and the lane fork/join.
static int queue_stream_create(cdet_ctx* ctx, hipStream_t* out) {
    const int words = (ctx->ncu + 31) / 32;
    std::vector<uint32_t> mask(static_cast<size_t>(words), 0u);
    for (int cu = 0; cu < ctx->ncu; ++cu) mask[static_cast<size_t>(cu / 32)] |= 1u << (cu % 32);
    HIP_TRY(hipExtStreamCreateWithCUMask(out, static_cast<uint32_t>(words), mask.data()));
    return 0;
}

#include "cdet_producer.hpp"
#include "cdet_plan.hpp"
#include "cdet_chunks.hpp"
#include "cdet_seed.hpp"

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* cdet_last_error(void) { return g_last_error.c_str(); }

int cdet_params_derive(cdet_params* p) {
    if (!p) return fail(CDET_ERR_ARGUMENT, "null params");
    if (!(p->U != 0.0)) return fail(CDET_ERR_ARGUMENT, "U must be nonzero");
    // sqrt(U*N1*N2*beta): exact double-double product, then one Newton
    // correction of the double square root, giving the correctly rounded
    // result (the reference rounds a 100-digit square root).
    double e1;
    const double p1 = p->U * p->beta;
    e1 = std::fma(p->U, p->beta, -p1);
    const double nn = static_cast<double>(p->N1) * static_cast<double>(p->N2);
    const double hi0 = p1 * nn;
    double e2 = std::fma(p1, nn, -hi0);
    const double lo0 = e2 + e1 * nn;
    double hi, lo;
    hi = dd_quick_two_sum(hi0, lo0, lo);
    const double s = std::sqrt(hi);
    const double r = std::fma(-s, s, hi) + lo;
    p->scale = s + r / (2.0 * s);
    p->diag_up_shift = p->alpha_down / p->U;
    p->diag_down_shift = p->alpha_up / p->U;
    return 0;
}

// Device state of a new context: stream, mask lists, counters, the first
// plain-path workspace, kernel attributes.
static int ctx_init(cdet_ctx* ctx) {
    if (hipGetDevice(&ctx->device) != hipSuccess) return fail(CDET_ERR_DEVICE, "no HIP device");
    if (hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess ||
        ctx->ncu <= 0)
        ctx->ncu = 256;
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    std::vector<uint16_t> masks;
    std::vector<int> off;
    masks_by_popcount(masks, off);
    HIP_TRY(hipMalloc(&ctx->d_masks, masks.size() * sizeof(uint16_t)));
    HIP_TRY(hipMalloc(&ctx->d_popoff, off.size() * sizeof(int)));
    HIP_TRY(hipMemcpy(ctx->d_masks, masks.data(), masks.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->d_popoff, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&ctx->d_fallbacks, sizeof(uint64_t)));
    HIP_TRY(hipMemset(ctx->d_fallbacks, 0, sizeof(uint64_t)));
    HIP_TRY(hipMalloc(&ctx->d_pair_cache, 2 * sizeof(unsigned long long)));
    HIP_TRY(hipMemset(ctx->d_pair_cache, 0, 2 * sizeof(unsigned long long)));
    HIP_TRY(hipMalloc(&ctx->d_one, sizeof(int32_t)));
    HIP_TRY(hipMalloc(&ctx->d_zero64, sizeof(int64_t)));
    HIP_TRY(hipMemset(ctx->d_one, 0, sizeof(int32_t)));
    HIP_TRY(hipMemset(ctx->d_zero64, 0, sizeof(int64_t)));
    HIP_TRY(hipEventCreateWithFlags(&ctx->query.done, hipEventDisableTiming));
    PlainWork* w = nullptr;
    int rc = plain_work_make(&w);
    if (rc) return rc;
    ctx->plain_all.push_back(w);
    ctx->plain_idle.push_back(w);
    // Allow the order-bucketed kernel the full LDS of a CU.
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_fpm<false, 256>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kMaxLds)));
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_fpm<false, 1024>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kMaxLds)));
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_cdet<kCdetLayLds, 256>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kCdetMaxDyn)));
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_cdet<kCdetLayLds, 1024>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kCdetMaxDyn)));
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_cdet<kCdetLaySplit, 256>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kCdetMaxDyn)));
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_cdet<kCdetLaySplit, 1024>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kCdetMaxDyn)));
    rc = noalloc_scratch_of(ctx->device, &ctx->noalloc_scratch);
    if (rc) return rc;
    ctx->tuning = default_tuning();
    refresh_scalars(ctx);
    return 0;
}

int cdet_ctx_create(const cdet_params* params, cdet_ctx** out_ctx) {
    if (!params || !out_ctx) return fail(CDET_ERR_ARGUMENT, "null argument");
    if (params->N1 <= 0 || params->N2 <= 0 || params->cheb_nu <= 0 || params->cheb_nv <= 0)
        return fail(CDET_ERR_ARGUMENT, "lattice extents and Chebyshev counts must be positive");
    if (params->cheb_nu > 256) return fail(CDET_ERR_CAPACITY, "cheb_nu above 256 is not supported");
    auto* ctx = new cdet_ctx();
    ctx->p = *params;
    const int rc = ctx_init(ctx);
    if (rc) {
        const std::string msg = g_last_error;
        cdet_ctx_destroy(ctx);
        return fail(rc, msg);
    }
    *out_ctx = ctx;
    return 0;
}

// Teardown in three phases: first everything that can still run on the
// device is stopped or waited for (slot threads joined, every stream and the
// last lookup call synchronised), then the plans the context made are
// destroyed, then the buffers are freed and the streams destroyed.  Plans a
// caller created with cdet_batch_plan_create must be destroyed before.
int cdet_ctx_destroy(cdet_ctx* ctx) {
    if (!ctx) return 0;
    for (auto& sl : ctx->slots) {
        if (sl.worker.joinable()) {
            {
                std::lock_guard<std::mutex> lk(sl.mu);
                sl.quit = true;
            }
            sl.cv.notify_all();
            sl.worker.join();
        }
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        for (hipStream_t l : sl.lanes)
            if (l) (void)hipStreamSynchronize(l);
    }
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (hipStream_t l : ctx->lanes)
        if (l) (void)hipStreamSynchronize(l);
    for (PlainWork* w : ctx->plain_all)
        if (w->stream) (void)hipStreamSynchronize(w->stream);
    if (ctx->query.done) (void)hipEventSynchronize(ctx->query.done);
    if (ctx->host_plan) cdet_batch_plan_destroy(ctx->host_plan);
    ctx->host_plan = nullptr;
    for (auto& sl : ctx->slots) {
        if (sl.plan) cdet_batch_plan_destroy(sl.plan);
        sl.plan = nullptr;
    }
    for (auto& t : ctx->tab) {
        if (t.alias < 0) {
            (void)hipFree(t.upper);
            (void)hipFree(t.lower);
            (void)hipFree(t.ucorner);
            (void)hipFree(t.lcorner);
            invalidate_split(t);
        }
    }
    (void)hipFree(ctx->d_masks);
    (void)hipFree(ctx->d_popoff);
    (void)hipFree(ctx->d_fallbacks);
    (void)hipFree(ctx->d_pair_cache);
    (void)hipFree(ctx->d_one);
    (void)hipFree(ctx->d_zero64);
    for (PlainWork* w : ctx->plain_all) plain_work_free(w);
    ctx->plain_all.clear();
    ctx->plain_idle.clear();
    query_scratch_free(ctx->query);
    if (ctx->query.done) (void)hipEventDestroy(ctx->query.done);
    for (auto& sl : ctx->slots) {
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        for (hipStream_t l : sl.lanes)
            if (l) (void)hipStreamDestroy(l);
        (void)hipFree(sl.d_v);
        (void)hipFree(sl.d_out);
        (void)hipFree(sl.d_status);
        (void)hipHostFree(sl.h_v);
        (void)hipHostFree(sl.h_out);
        (void)hipHostFree(sl.h_status);
    }
    for (hipStream_t l : ctx->lanes)
        if (l) (void)hipStreamDestroy(l);
    (void)hipHostFree(ctx->h_hv);
    (void)hipHostFree(ctx->h_hout);
    for (auto& hb : ctx->host_bufs) (void)hipHostFree(const_cast<char*>(hb.first));
    (void)hipHostFree(ctx->h_hstatus);
    (void)hipFree(ctx->d_hv);
    (void)hipFree(ctx->d_hout);
    (void)hipFree(ctx->d_hstatus);
    (void)hipFree(ctx->d_seed_tab);
    (void)hipFree(ctx->d_seed_in);
    (void)hipFree(ctx->d_seed_out);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

void* cdet_ctx_stream(cdet_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int cdet_stream_create(cdet_ctx* ctx, void** out) {
    if (!ctx || !out) return fail(CDET_ERR_ARGUMENT, "cdet_stream_create: bad argument");
    *out = nullptr;
    hipStream_t s = nullptr;
    const int rc = queue_stream_create(ctx, &s);
    if (rc) return rc;
    *out = s;
    return 0;
}

int cdet_stream_destroy(cdet_ctx* ctx, void* stream) {
    if (!ctx || !stream) return fail(CDET_ERR_ARGUMENT, "cdet_stream_destroy: bad argument");
    HIP_TRY(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    return 0;
}

static int table_sizes(const cdet_ctx* ctx, size_t& ncoef, size_t& ncorner) {
    const size_t S = static_cast<size_t>(ctx->p.N1) * ctx->p.N2;
    ncorner = 4 * S * S;
    ncoef = ncorner * static_cast<size_t>(ctx->p.cheb_nu) * ctx->p.cheb_nv;
    return 0;
}

// Drop the copies (FP32 split, step order) derived from a table whose values
// change.
static void invalidate_split(DevTable& t) {
    for (auto& b : t.split)
        for (float*& f : b) {
            (void)hipFree(f);
            f = nullptr;
        }
    for (double*& d : t.steps) {
        (void)hipFree(d);
        d = nullptr;
    }
    for (double*& d : t.abssum) {
        (void)hipFree(d);
        d = nullptr;
    }
    (void)hipFree(t.rowsum);
    t.rowsum = nullptr;
}

static int alloc_table(cdet_ctx* ctx, int kind) {
    DevTable& t = ctx->tab[kind];
    if (t.alias >= 0) {
        t = DevTable{};
    }
    if (t.upper) return 0;
    size_t nc, nk;
    table_sizes(ctx, nc, nk);
    HIP_TRY(hipMalloc(&t.upper, nc * sizeof(double)));
    HIP_TRY(hipMalloc(&t.lower, nc * sizeof(double)));
    HIP_TRY(hipMalloc(&t.ucorner, nk * sizeof(double)));
    HIP_TRY(hipMalloc(&t.lcorner, nk * sizeof(double)));
    return 0;
}

int cdet_tables_upload(cdet_ctx* ctx, int kind, const double* upper, const double* lower, const double* ucorner,
                       const double* lcorner) {
    if (!ctx || kind < 0 || kind > 3 || !upper || !lower || !ucorner || !lcorner)
        return fail(CDET_ERR_ARGUMENT, "cdet_tables_upload: bad argument");
    for (int k = 0; k < 4; ++k)
        if (k != kind && ctx->tab[k].alias == kind) return fail(CDET_ERR_TABLES, "table kind is aliased by another kind");
    int rc = alloc_table(ctx, kind);
    if (rc) return rc;
    size_t nc, nk;
    table_sizes(ctx, nc, nk);
    DevTable& t = ctx->tab[kind];
    invalidate_split(t);
    if (kind >= CDET_TABLE_ORDINARY_UP) ctx->ordinary_layout = CDET_ORDINARY_RESIDENT;
    HIP_TRY(hipMemcpy(t.upper, upper, nc * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.lower, lower, nc * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.ucorner, ucorner, nk * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.lcorner, lcorner, nk * sizeof(double), hipMemcpyHostToDevice));
    refresh_scalars(ctx);
    return 0;
}

int cdet_tables_alias(cdet_ctx* ctx, int dst, int src) {
    if (!ctx || dst < 0 || dst > 3 || src < 0 || src > 3 || dst == src)
        return fail(CDET_ERR_ARGUMENT, "cdet_tables_alias: bad kinds");
    if (ctx->tab[src].alias >= 0) src = ctx->tab[src].alias;
    if (!ctx->tab[src].upper) return fail(CDET_ERR_TABLES, "cdet_tables_alias: source kind has no table");
    DevTable& t = ctx->tab[dst];
    if (t.alias < 0 && t.upper) {
        (void)hipFree(t.upper);
        (void)hipFree(t.lower);
        (void)hipFree(t.ucorner);
        (void)hipFree(t.lcorner);
        invalidate_split(t);
    }
    t = DevTable{};
    t.alias = src;
    refresh_scalars(ctx);
    return 0;
}

int cdet_tables_download(cdet_ctx* ctx, int kind, double* upper, double* lower, double* ucorner, double* lcorner) {
    if (!ctx || kind < 0 || kind > 3) return fail(CDET_ERR_ARGUMENT, "cdet_tables_download: bad argument");
    const int k = ctx->tab[kind].alias >= 0 ? ctx->tab[kind].alias : kind;
    const DevTable& t = ctx->tab[k];
    if (!t.upper) return fail(CDET_ERR_TABLES, "cdet_tables_download: kind has no table");
    size_t nc, nk;
    table_sizes(ctx, nc, nk);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (upper) HIP_TRY(hipMemcpy(upper, t.upper, nc * sizeof(double), hipMemcpyDeviceToHost));
    if (lower) HIP_TRY(hipMemcpy(lower, t.lower, nc * sizeof(double), hipMemcpyDeviceToHost));
    if (ucorner) HIP_TRY(hipMemcpy(ucorner, t.ucorner, nk * sizeof(double), hipMemcpyDeviceToHost));
    if (lcorner) HIP_TRY(hipMemcpy(lcorner, t.lcorner, nk * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int cdet_tables_download_blocks(cdet_ctx* ctx, int kind, const int64_t* sitepairs, int64_t nblocks, double* upper,
                                double* lower, double* ucorner, double* lcorner) {
    if (!ctx || kind < 0 || kind > 3 || nblocks < 0 || (nblocks > 0 && !sitepairs))
        return fail(CDET_ERR_ARGUMENT, "cdet_tables_download_blocks: bad argument");
    const int k = ctx->tab[kind].alias >= 0 ? ctx->tab[kind].alias : kind;
    const DevTable& t = ctx->tab[k];
    if (!t.upper) return fail(CDET_ERR_TABLES, "cdet_tables_download_blocks: kind has no table");
    size_t nc, nk;
    table_sizes(ctx, nc, nk);
    const size_t blk = static_cast<size_t>(ctx->p.cheb_nu) * ctx->p.cheb_nv;
    for (int64_t i = 0; i < nblocks; ++i)
        if (sitepairs[i] < 0 || static_cast<size_t>(sitepairs[i]) >= nk)
            return fail(CDET_ERR_ARGUMENT, "cdet_tables_download_blocks: sitepair out of range");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < nblocks; ++i) {
        const size_t sp = static_cast<size_t>(sitepairs[i]);
        if (upper) HIP_TRY(hipMemcpy(upper + i * blk, t.upper + sp * blk, blk * sizeof(double), hipMemcpyDeviceToHost));
        if (lower) HIP_TRY(hipMemcpy(lower + i * blk, t.lower + sp * blk, blk * sizeof(double), hipMemcpyDeviceToHost));
        if (ucorner) HIP_TRY(hipMemcpy(ucorner + i, t.ucorner + sp, sizeof(double), hipMemcpyDeviceToHost));
        if (lcorner) HIP_TRY(hipMemcpy(lcorner + i, t.lcorner + sp, sizeof(double), hipMemcpyDeviceToHost));
    }
    return 0;
}

// One configuration on a workspace of the context's pool: host-to-device
// copy, launches, device-to-host copy, wait, all on the workspace's stream.
// Threads that call at the same time hold different workspaces.
static int plain_run(cdet_ctx* ctx, PlainWork* w, const cdet_vertex* v, int n, int mode, double* out,
                     size_t out_count) {
    hipStream_t s = w->stream;
    HIP_TRY(hipMemcpyAsync(w->d_verts, v, sizeof(double) * 4 * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, w->d_small_i32, static_cast<int64_t>(8), 0);
    hipLaunchKernelGGL(k_pairs_plain, dim3(2 * n * n), dim3(64), 0, s, ctx->ks, w->d_verts, n, w->d_gup, w->d_gdn);
    HIP_TRY(hipGetLastError());
    MinorsOut o{w->d_out, w->d_out_ord, w->d_single, w->d_small_i32, w->d_small_i32 + 2, w->d_small_i32 + 1};
    int rc = launch_minors(ctx, w->scratch, n, mode, ctx->d_one, 1, ctx->d_zero64, ctx->d_zero64, w->d_gup, w->d_gdn,
                           o, s);
    if (rc) return rc;
    rc = launch_combine(mode, 1, static_cast<int64_t>(out_count), ctx->d_zero64, o, s);
    if (rc) return rc;
    int32_t st = 0;
    HIP_TRY(hipMemcpyAsync(out, w->d_out, sizeof(double) * out_count, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&st, w->d_small_i32, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    ctx->calls += 1;
    if ((st & 1) && n > 1) {
        if (mode == 1) return fail(CDET_ERR_NONFINITE, "Error (integrand): Value is not finite");
        return fail(CDET_ERR_NONFINITE, "Error (all-subset integrand): Value is not finite");
    }
    return 0;
}

static int plain_eval(cdet_ctx* ctx, const cdet_vertex* v, int n, int mode, double* out, size_t out_count) {
    if (!tables_ready(ctx)) return fail(CDET_ERR_TABLES, "Resident Green-function tables have not been uploaded");
    PlainWork* w = nullptr;
    int rc = plain_work_acquire(ctx, &w);
    if (rc) return rc;
    rc = plain_run(ctx, w, v, n, mode, out, out_count);
    if (rc) (void)hipStreamSynchronize(w->stream);  // nothing of a failed call stays in flight
    plain_work_release(ctx, w);
    return rc;
}

int cdet_eval_all_subsets(cdet_ctx* ctx, const cdet_vertex* vertices, int envelope_order, double* out_values) {
    if (!ctx || !vertices || !out_values) return fail(CDET_ERR_ARGUMENT, "null argument");
    if (envelope_order <= 1)
        return fail(CDET_ERR_ORDER, "The all-subset replica integrand is defined only for envelope order >= 2");
    if (envelope_order > kMaxOrder) return fail(CDET_ERR_CAPACITY, "envelope order above CDET_MAX_ORDER");
    return plain_eval(ctx, vertices, envelope_order, 0, out_values, size_t(1) << envelope_order);
}

int cdet_eval_single(cdet_ctx* ctx, const cdet_vertex* vertices, int n, double* out_value) {
    if (!ctx || !vertices || !out_value) return fail(CDET_ERR_ARGUMENT, "null argument");
    if (n <= 0) return fail(CDET_ERR_ORDER, "The replica integrand is defined only for order >= 1");
    if (n > kMaxOrder) return fail(CDET_ERR_CAPACITY, "order above CDET_MAX_ORDER");
    return plain_eval(ctx, vertices, n, 1, out_value, 1);
}

// Layout of the no-allocation scratch for one configuration of order n.
struct NoAllocLayout {
    size_t gup, gdn, out_ord, single, small, minors, minors_cap, total;
};
static NoAllocLayout noalloc_layout(int n) {
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    NoAllocLayout L{};
    size_t o = 0;
    L.gup = o;
    o += up(sizeof(double) * 2 * static_cast<size_t>(n) * n);
    L.gdn = o;
    o += up(sizeof(double) * 2 * static_cast<size_t>(n) * n);
    L.out_ord = o;
    o += up(sizeof(double) << n);
    L.single = o;
    o += up(sizeof(double) * 4);
    L.small = o;
    o += up(sizeof(int32_t) * 8);
    L.minors = o;
    L.minors_cap = minors_scratch_bytes(n, 1, 0);
    o += up(L.minors_cap);
    L.total = o;
    return L;
}

int cdet_eval_all_subsets_noalloc(cdet_ctx* ctx, const cdet_vertex* d_vertices, int envelope_order,
                                  double* d_out, int32_t* d_status, void* stream) {
    if (!ctx || !d_vertices || !d_out) return fail(CDET_ERR_ARGUMENT, "null argument");
    if (envelope_order <= 1)
        return fail(CDET_ERR_ORDER, "The all-subset replica integrand is defined only for envelope order >= 2");
    if (envelope_order > kMaxOrder) return fail(CDET_ERR_CAPACITY, "envelope order above CDET_MAX_ORDER");
    if (!tables_ready(ctx)) return fail(CDET_ERR_TABLES, "Resident Green-function tables have not been uploaded");
    const int n = envelope_order;
    const NoAllocLayout L = noalloc_layout(n);
    if (L.total > kNoAllocBytes)
        return fail(CDET_ERR_CAPACITY, "cdet_eval_all_subsets_noalloc: order " + std::to_string(n) + " needs " +
                                           std::to_string(L.total) + " bytes of scratch, above the " +
                                           std::to_string(kNoAllocBytes) + "-byte static buffer");
    unsigned char* base = ctx->noalloc_scratch;
    if (!base) return fail(CDET_ERR_DEVICE, "no-allocation scratch was not set up");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    double* gup = reinterpret_cast<double*>(base + L.gup);
    double* gdn = reinterpret_cast<double*>(base + L.gdn);
    int32_t* small = reinterpret_cast<int32_t*>(base + L.small);
    int32_t* status = d_status ? d_status : small;
    // status word and fallback queue start at zero (fill kernels, no memset)
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, small, static_cast<int64_t>(8), 0);
    if (d_status) hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, d_status, static_cast<int64_t>(1), 0);
    hipLaunchKernelGGL(k_pairs_plain, dim3(2 * n * n), dim3(64), 0, s, ctx->ks,
                       reinterpret_cast<const double*>(d_vertices), n, gup, gdn);
    HIP_TRY(hipGetLastError());
    MinorsOut o{d_out, reinterpret_cast<double*>(base + L.out_ord), reinterpret_cast<double*>(base + L.single),
                status, small + 2, small + 1};
    // the minors scratch is a fixed region: its capacity covers this order,
    // so launch_minors finds it large enough and allocates nothing
    Scratch region;
    region.ptr = base + L.minors;
    region.cap = L.minors_cap;
    int rc = launch_minors(ctx, region, n, 0, ctx->d_one, 1, ctx->d_zero64, ctx->d_zero64, gup, gdn, o, s);
    if (rc) return rc;
    rc = launch_combine(0, 1, static_cast<int64_t>(size_t(1) << n), ctx->d_zero64, o, s);
    if (rc) return rc;
    ctx->calls += 1;
    return 0;
}

int cdet_ctx_set_tuning(cdet_ctx* ctx, const cdet_tuning* in) {
    if (!ctx || !in) return fail(CDET_ERR_ARGUMENT, "null argument");
    cdet_tuning t = default_tuning();
    if (in->contract_wg_per_cu > 0) {
        if (in->contract_wg_per_cu > 8) return fail(CDET_ERR_ARGUMENT, "contract_wg_per_cu must be 1..8");
        t.contract_wg_per_cu = in->contract_wg_per_cu;
    }
    if (in->contract_tile > 0) {
        if (in->contract_tile != 16 && in->contract_tile != kF64Tile)
            return fail(CDET_ERR_ARGUMENT, "contract_tile must be 16 or 64");
        t.contract_tile = in->contract_tile;
    }
    if (in->pair_min_per_cu > 0) t.pair_min_per_cu = in->pair_min_per_cu;
    if (in->max_lanes > 0) {
        if (in->max_lanes > kMinorLanes) return fail(CDET_ERR_ARGUMENT, "max_lanes must be 1..4");
        t.max_lanes = in->max_lanes;
    }
    for (int n = 0; n <= kMaxOrder; ++n) {
        if (in->minors_threads[n] > 0) {
            if (in->minors_threads[n] != 256 && in->minors_threads[n] != 1024)
                return fail(CDET_ERR_ARGUMENT, "minors_threads must be 256 or 1024");
            t.minors_threads[n] = in->minors_threads[n];
        }
        if (in->group16_min_terms[n] > 0) t.group16_min_terms[n] = std::max(16, in->group16_min_terms[n]);
        if (in->group64_min_terms[n] > 0) t.group64_min_terms[n] = std::max(64, in->group64_min_terms[n]);
    }
    if (in->fused_minors_max_order > 0) {
        if (in->fused_minors_max_order > kPairOrderMax)
            return fail(CDET_ERR_ARGUMENT, "fused_minors_max_order must be 1..10");
        t.fused_minors_max_order = in->fused_minors_max_order;
    }
    ctx->tuning = t;
    refresh_scalars(ctx);  // plans made afterwards capture the new thresholds
    return 0;
}

int cdet_ctx_set_sum_order(cdet_ctx* ctx, int order) {
    if (!ctx) return fail(CDET_ERR_ARGUMENT, "null argument");
    if (order != CDET_SUM_TREE && order != CDET_SUM_REFERENCE)
        return fail(CDET_ERR_ARGUMENT, "sum order must be CDET_SUM_TREE or CDET_SUM_REFERENCE");
    ctx->sum_order = order;
    refresh_scalars(ctx);  // a new generation: captured plan graphs are captured again
    return 0;
}

int cdet_ctx_get_tuning(cdet_ctx* ctx, cdet_tuning* out) {
    if (!ctx || !out) return fail(CDET_ERR_ARGUMENT, "null argument");
    *out = ctx->tuning;
    return 0;
}

int cdet_get_counters(cdet_ctx* ctx, cdet_counters* out) {
    if (!ctx || !out) return fail(CDET_ERR_ARGUMENT, "null argument");
    uint64_t fb = 0;
    unsigned long long pc[2] = {0, 0};
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemcpy(&fb, ctx->d_fallbacks, sizeof(uint64_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(pc, ctx->d_pair_cache, sizeof(pc), hipMemcpyDeviceToHost));
    out->integrand_calls = ctx->calls;
    out->fpm_fallbacks = fb;
    out->green_pair_cache_hits = pc[0];
    out->green_pair_cache_misses = pc[1];
    return 0;
}

int cdet_reset_counters(cdet_ctx* ctx) {
    if (!ctx) return fail(CDET_ERR_ARGUMENT, "null argument");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemset(ctx->d_fallbacks, 0, sizeof(uint64_t)));
    HIP_TRY(hipMemset(ctx->d_pair_cache, 0, 2 * sizeof(unsigned long long)));
    ctx->calls = 0;
    return 0;
}

}  // extern "C"

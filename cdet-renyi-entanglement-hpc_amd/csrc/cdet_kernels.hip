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
//   k_fpm_cdet<n>    order-bucketed: one workgroup per configuration, vertex
//                    matrices, principal minors and subset recursion in LDS
// The plain form (one configuration per launch) uses k_pairs_plain (one wave
// per Green pair) and the same k_fpm_cdet.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cdet_hip.h"
#include "../../include/cdet_mcmcmc.h"
#include "cdet_cdet.hpp"
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
__global__ __launch_bounds__(256) void k_plan(KScalars ks, const double* __restrict__ verts,
                                              const int32_t* __restrict__ cfg_order,
                                              const int64_t* __restrict__ cfg_vbase,
                                              const int64_t* __restrict__ cfg_pbase, double* __restrict__ g_up,
                                              double* __restrict__ g_dn, int32_t* __restrict__ q_key,
                                              int32_t* __restrict__ q_rank, double* __restrict__ q_u,
                                              double* __restrict__ q_v, int32_t* __restrict__ hist) {
    __shared__ Vtx vs[2][kMaxOrder];
    const int c = blockIdx.x;
    const int n = cfg_order[c];
    const int64_t vb = cfg_vbase[c];
    const int64_t qb = 2 * cfg_pbase[c];
    if (threadIdx.x < 2 * n) {
        const int cls = threadIdx.x / n;
        const int v = threadIdx.x - cls * n;
        vs[cls][v] = decode_vertex(verts + 4 * (vb + v), ks.beta, ks.N1, ks.N2, cls == 0 ? 2 : 1);
    }
    __syncthreads();
    const int n2 = n * n;
    for (int t = threadIdx.x; t < 2 * n2; t += blockDim.x) {
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
        if (d.corner) {
            g_up[q] = ks.corner[ku][br][d.sitepair];
            g_dn[q] = ks.corner[ku + 1][br][d.sitepair];
            q_key[q] = -1;
            continue;
        }
        const int32_t key = static_cast<int32_t>((static_cast<int64_t>(cls) * 2 + br) * ks.SP + d.sitepair);
        q_key[q] = key;
        q_rank[q] = atomicAdd(&hist[key], 1);
        q_u[q] = d.u;
        q_v[q] = d.v;
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
    q_key[q] = key;
    q_rank[q] = atomicAdd(&hist[key], 1);
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
__global__ __launch_bounds__(1024) void k_scan_keys(const int32_t* __restrict__ hist, int32_t nkeys, int32_t tile,
                                                    int32_t* __restrict__ key_off, int32_t* __restrict__ tile_key,
                                                    int32_t* __restrict__ tile_start, int32_t* __restrict__ tile_len,
                                                    int32_t* __restrict__ n_tiles) {
    __shared__ int32_t s_q[1024];
    __shared__ int32_t s_t[1024];
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const int chunk = (nkeys + nt - 1) / nt;
    const int k0 = min(nkeys, tid * chunk);
    const int k1 = min(nkeys, k0 + chunk);
    int sq = 0, st = 0;
    for (int k = k0; k < k1; ++k) {
        const int h = hist[k];
        sq += h;
        st += (h + tile - 1) / tile;
    }
    s_q[tid] = sq;
    s_t[tid] = st;
    __syncthreads();
    for (int off = 1; off < nt; off <<= 1) {
        int aq = 0, at = 0;
        if (tid >= off) {
            aq = s_q[tid - off];
            at = s_t[tid - off];
        }
        __syncthreads();
        s_q[tid] += aq;
        s_t[tid] += at;
        __syncthreads();
    }
    int oq = s_q[tid] - sq;
    int ot = s_t[tid] - st;
    for (int k = k0; k < k1; ++k) {
        const int h = hist[k];
        key_off[k] = oq;
        for (int s = 0; s < h; s += tile) {
            tile_key[ot] = k;
            tile_start[ot] = oq + s;
            tile_len[ot] = min(tile, h - s);
            ++ot;
        }
        oq += h;
    }
    if (tid == nt - 1) *n_tiles = s_t[tid];
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

// Exact batched contraction: one wave per tile of <= 64 queries sharing a
// coefficient block; lane = query.  The block address is wave-uniform, so the
// coefficients stream through the scalar cache and are broadcast to all
// lanes.  Rows are produced from a = nu-1 down to 0 and consumed at once by
// the Clenshaw over u, so no row buffer is needed.  Bit-identical to the
// reference's chebyshev_sum_2d_pair_from_ptrs (include/cdet/g_funcs.h:687-763).
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
    const int total = *n_tiles;
    for (int t = blockIdx.x * waves_per_block + wave; t < total; t += gridDim.x * waves_per_block) {
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
        const double two_v = 2.0 * v;
        UClenshaw fu, fd;
        fu.init(u, ks.nu);
        fd.init(u, ks.nu);
        double gu = 0.0, gd = 0.0;
        const int dv = ks.nv - 1;
        for (int a = ks.nu - 1; a >= 0; --a) {
            const double* ru = cu + static_cast<int64_t>(a) * ks.nv;
            double r_u;
            if (dv >= 2) {
                double b1 = 0.0, b2 = 0.0;
                for (int k = dv; k >= 1; --k) {
                    const double bk = ru[k] + two_v * b1 - b2;
                    b2 = b1;
                    b1 = bk;
                }
                r_u = ru[0] + b1 * v - b2;
            } else {
                r_u = clenshaw_row(ru, dv, v);
            }
            gu = fu.feed(a, r_u);
            if (!same) {
                const double* rd = cd + static_cast<int64_t>(a) * ks.nv;
                double r_d;
                if (dv >= 2) {
                    double b1 = 0.0, b2 = 0.0;
                    for (int k = dv; k >= 1; --k) {
                        const double bk = rd[k] + two_v * b1 - b2;
                        b2 = b1;
                        b1 = bk;
                    }
                    r_d = rd[0] + b1 * v - b2;
                } else {
                    r_d = clenshaw_row(rd, dv, v);
                }
                gd = fd.feed(a, r_d);
            }
        }
        if (active) {
            const int32_t dest = s_dest[pos];
            g_up[dest] = gu;
            g_dn[dest] = same ? gu : gd;
        }
    }
}

// Order-bucketed vertex-matrix -> minors -> connected values -> assembly.
// One workgroup per configuration of order NORD; LDS holds everything.
//   mode 0: all-subset values into out[cfg_obase[c] .. + 2^n)
//   mode 1: single-order value into out[cfg_obase[c]]
// Configurations are taken from cfg_list (indices into the plan arrays).
struct FpmLds {
    static __host__ __device__ inline size_t bytes(int n) {
        const size_t sc = size_t(1) << n;
        const size_t list = static_cast<size_t>(fpm_list_size(n));
        const size_t inv = n >= 2 ? (size_t(1) << (n - 2)) : 1;
        return sizeof(double) * (2 * n * n + 2 * n + 2 * sc + 2 * list + inv + 256) + 64;
    }
};

__global__ __launch_bounds__(256) void k_fpm_cdet(KScalars ks, int n, int mode, const int32_t* __restrict__ cfg_list,
                                                  const int64_t* __restrict__ cfg_pbase,
                                                  const int64_t* __restrict__ cfg_obase, const double* __restrict__ g_up,
                                                  const double* __restrict__ g_dn, double* __restrict__ out,
                                                  int32_t* __restrict__ status, uint64_t* __restrict__ fallback_count,
                                                  const uint16_t* __restrict__ masks_by_pop,
                                                  const int* __restrict__ pop_off_all, double* __restrict__ gscratch,
                                                  size_t gscratch_stride) {
    extern __shared__ double lds_dyn[];
    __shared__ int s_flag;
    __shared__ int s_status;
    __shared__ double s_single[2][2];
    const int c = cfg_list[blockIdx.x];
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const uint32_t sc = 1u << n;
    const int n2 = n * n;
    const int list = fpm_list_size(n);
    const int ninv = n >= 2 ? (1 << (n - 2)) : 1;
    double* base = gscratch ? gscratch + static_cast<size_t>(blockIdx.x) * gscratch_stride : lds_dyn;
    double* mU = base;
    double* mD = mU + n2;
    double* dU = mD + n2;
    double* dD = dU + n;
    double* smU = dD + n;
    double* smD = smU + sc;
    double* bufA = smD + sc;
    double* bufB = bufA + list;
    double* inv = bufB + list;
    double* red = inv + ninv;
    const int* pop_off = pop_off_all + n * (kMaxOrder + 2);
    const int64_t pb = 2 * cfg_pbase[c];
    const int64_t ob = cfg_obase[c];
    if (tid == 0) s_status = 0;
    __syncthreads();

    for (int cls = 0; cls < 2; ++cls) {
        const bool same = ks.same_spin[cls] != 0;
        const double* gu = g_up + pb + cls * n2;
        const double* gd = g_dn + pb + cls * n2;
        // Vertex matrices (include/cdet/integrand_replica_randomly_sampled.h:751-761).
        for (int idx = tid; idx < n2; idx += nt) {
            const int i = idx / n;
            const int j = idx - i * n;
            double vu = gu[idx];
            double vd = gd[idx];
            if (i == j) {
                dU[i] = vu + ks.dup;
                dD[i] = vd + ks.ddn;
                vu += ks.dup;
                vd += ks.ddn;
            }
            mU[idx] = vu * ks.scale;
            mD[idx] = vd * ks.scale;
        }
        __syncthreads();
        if (n == 1) {
            // Order-1 single value (include/cdet/integrand_replica_randomly_sampled.h:173-213).
            if (tid == 0) {
                const double v = ks.c1[cls] * (dU[0] * dD[0] - ks.counter);
                s_single[cls][0] = v;
                s_single[cls][1] = 0.0;
            }
            __syncthreads();
            continue;
        }
        int fb = fpm_block(mU, n, ks.tol, ks.shift, smU, bufA, bufB, inv, red, &s_flag);
        if (same) {
            for (uint32_t m = tid; m < sc; m += nt) smD[m] = smU[m];
            __syncthreads();
            fb *= 2;
        } else {
            fb += fpm_block(mD, n, ks.tol, ks.shift, smD, bufA, bufB, inv, red, &s_flag);
        }
        if (fb && tid == 0) {
            atomicAdd(reinterpret_cast<unsigned long long*>(fallback_count), static_cast<unsigned long long>(fb));
            s_status |= 2;
        }
        // a[mask] = pm_up * pm_down with a[0] = 0 (stored over smU).
        for (uint32_t m = tid; m < sc; m += nt) smU[m] = (m == 0) ? 0.0 : smU[m] * smD[m];
        __syncthreads();
        if (mode == 0) {
            ztransform_all_block(smU, smD, n, ks.thr, ks.filter != 0, masks_by_pop, pop_off);
            // Assembly (:833-877) and twisted - 2*ordinary
            // (include/cdet/cdet_integrand_evaluator.h:348-350).  Thread t
            // owns the same masks in both classes.
            const int rc = cls == 0 ? 2 : 1;
            (void)rc;
            bool bad = false;
            for (uint32_t m = tid; m < sc; m += nt) {
                double v;
                if (m == 0) {
                    v = 0.0;
                } else if ((m & (m - 1)) == 0) {
                    const int i = __builtin_ctz(m);
                    v = ks.c1[cls] * (dU[i] * dD[i] - ks.counter);
                } else {
                    v = ks.pref[cls][__builtin_popcount(m)] * smD[m];
                }
                if (!isfinite(v)) bad = true;
                if (cls == 0)
                    out[ob + m] = v;
                else
                    out[ob + m] = out[ob + m] - 2.0 * v;
            }
            if (bad) atomicOr(&s_status, 1);
            __syncthreads();
        } else {
            const double cres = ztransform_single_block(smU, smD, n, ks.thr, ks.filter != 0, masks_by_pop, pop_off);
            if (tid == 0) {
                const DD v = dd_mul(DD{ks.pref_hi[cls][n], ks.pref_lo[cls][n]}, DD{cres, 0.0});
                s_single[cls][0] = v.hi;
                s_single[cls][1] = v.lo;
            }
            __syncthreads();
        }
    }
    if (mode == 1 && tid == 0) {
        // twisted - 2*ordinary in double-double, then rounded
        // (include/cdet/cdet_integrand_evaluator.h:262, :559).
        const DD tw{s_single[0][0], s_single[0][1]};
        const DD od{s_single[1][0], s_single[1][1]};
        const DD r = dd_sub(tw, dd_mul(DD{2.0, 0.0}, od));
        const double v = r.hi + r.lo;
        if (!isfinite(tw.hi + tw.lo) || !isfinite(od.hi + od.lo) || !isfinite(v)) s_status |= 1;
        out[ob] = v;
    }
    __syncthreads();
    if (tid == 0 && status) status[c] = s_status;
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

// Fill kernel (graph-capture safe replacement of hipMemsetAsync).
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct cdet_batch_plan;

struct cdet_ctx {
    cdet_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    DevTable tab[4];
    KScalars ks{};
    uint16_t* d_masks = nullptr;  // masks by popcount for every order, (kMaxOrder+1) * 2^kMaxOrder
    int* d_popoff = nullptr;      // (kMaxOrder+1) * (kMaxOrder+2)
    uint64_t* d_fallbacks = nullptr;
    uint64_t calls = 0;
    // plain-path scratch
    double* d_verts = nullptr;
    double* d_gup = nullptr;
    double* d_gdn = nullptr;
    double* d_out = nullptr;
    int32_t* d_status = nullptr;
    int32_t* d_one = nullptr;   // cfg list {0}
    int64_t* d_zero64 = nullptr;  // pbase/obase {0}
    double* d_gscratch = nullptr;
    size_t gscratch_bytes = 0;
    // host-buffer batch path (driver callback)
    int batch_mode = CDET_MODE_EXACT;
    cdet_batch_plan* host_plan = nullptr;
    std::vector<int32_t> host_plan_orders;
    void* d_hv = nullptr;
    size_t d_hv_cap = 0;
    void* d_hout = nullptr;
    size_t d_hout_cap = 0;
    void* d_hstatus = nullptr;
    size_t d_hstatus_cap = 0;
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
        k.split_hi[kind][0] = t.split[0][0];
        k.split_lo[kind][0] = t.split[0][1];
        k.split_hi[kind][1] = t.split[1][0];
        k.split_lo[kind][1] = t.split[1][1];
    }
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

}  // namespace

// Sizing helper: LDS bytes of k_fpm_cdet for order n.
static size_t fpm_lds_bytes(int n) { return FpmLds::bytes(n); }
static constexpr size_t kMaxLds = 160 * 1024 - 2048;

// Launch the per-configuration kernel for one order bucket.
static int launch_fpm(cdet_ctx* ctx, int n, int mode, const int32_t* cfg_list, int count, const int64_t* pbase,
                      const int64_t* obase, const double* gup, const double* gdn, double* out, int32_t* status,
                      hipStream_t stream) {
    if (count <= 0) return 0;
    const size_t lds = fpm_lds_bytes(n);
    double* gs = nullptr;
    size_t stride = 0;
    size_t dyn = lds;
    if (lds > kMaxLds) {
        stride = lds / sizeof(double);
        const size_t need = stride * sizeof(double) * static_cast<size_t>(count);
        void* p = ctx->d_gscratch;
        int rc = grow(&p, ctx->gscratch_bytes, need);
        ctx->d_gscratch = static_cast<double*>(p);
        if (rc) return rc;
        gs = ctx->d_gscratch;
        dyn = 0;
    }
    hipLaunchKernelGGL(k_fpm_cdet, dim3(count), dim3(256), dyn, stream, ctx->ks, n, mode, cfg_list, pbase, obase, gup,
                       gdn, out, status, ctx->d_fallbacks, ctx->d_masks + static_cast<size_t>(n) * (1u << kMaxOrder),
                       ctx->d_popoff, gs, stride);
    HIP_TRY(hipGetLastError());
    return 0;
}

static int mfma_tile_queries(const KScalars&) { return kMfmaTile; }

// FP32 hi/lo copies of every resident table (split-FP32 mode), built once.
static int ensure_split_tables(cdet_ctx* ctx) {
    size_t nc, nk;
    const size_t S = static_cast<size_t>(ctx->p.N1) * ctx->p.N2;
    nk = 4 * S * S;
    nc = nk * static_cast<size_t>(ctx->p.cheb_nu) * ctx->p.cheb_nv;
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
        }
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    refresh_scalars(ctx);
    return 0;
}

static int contract_launch_mfma(const KScalars& ks, int mode, int grid, const int32_t* ntiles, const int32_t* tkey,
                                const int32_t* tstart, const int32_t* tlen, const double* su, const double* sv,
                                const int32_t* sdest, double* gup, double* gdn, hipStream_t s) {
    if (mode == CDET_MODE_MFMA64) {
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

static int table_sizes(const cdet_ctx* ctx, size_t& ncoef, size_t& ncorner);
static int alloc_table(cdet_ctx* ctx, int kind);

#include "cdet_producer.hpp"
#include "cdet_plan.hpp"

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

int cdet_ctx_create(const cdet_params* params, cdet_ctx** out_ctx) {
    if (!params || !out_ctx) return fail(CDET_ERR_ARGUMENT, "null argument");
    if (params->N1 <= 0 || params->N2 <= 0 || params->cheb_nu <= 0 || params->cheb_nv <= 0)
        return fail(CDET_ERR_ARGUMENT, "lattice extents and Chebyshev counts must be positive");
    if (params->cheb_nu > 256) return fail(CDET_ERR_CAPACITY, "cheb_nu above 256 is not supported");
    auto* ctx = new cdet_ctx();
    ctx->p = *params;
    if (hipGetDevice(&ctx->device) != hipSuccess) {
        delete ctx;
        return fail(CDET_ERR_DEVICE, "no HIP device");
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return fail(CDET_ERR_DEVICE, "hipStreamCreate failed");
    }
    std::vector<uint16_t> masks;
    std::vector<int> off;
    masks_by_popcount(masks, off);
    HIP_TRY(hipMalloc(&ctx->d_masks, masks.size() * sizeof(uint16_t)));
    HIP_TRY(hipMalloc(&ctx->d_popoff, off.size() * sizeof(int)));
    HIP_TRY(hipMemcpy(ctx->d_masks, masks.data(), masks.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->d_popoff, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&ctx->d_fallbacks, sizeof(uint64_t)));
    HIP_TRY(hipMemset(ctx->d_fallbacks, 0, sizeof(uint64_t)));
    const int Mx = kMaxOrder;
    HIP_TRY(hipMalloc(&ctx->d_verts, sizeof(double) * 4 * Mx));
    HIP_TRY(hipMalloc(&ctx->d_gup, sizeof(double) * 2 * Mx * Mx));
    HIP_TRY(hipMalloc(&ctx->d_gdn, sizeof(double) * 2 * Mx * Mx));
    HIP_TRY(hipMalloc(&ctx->d_out, sizeof(double) << Mx));
    HIP_TRY(hipMalloc(&ctx->d_status, sizeof(int32_t)));
    HIP_TRY(hipMalloc(&ctx->d_one, sizeof(int32_t)));
    HIP_TRY(hipMalloc(&ctx->d_zero64, sizeof(int64_t)));
    HIP_TRY(hipMemset(ctx->d_one, 0, sizeof(int32_t)));
    HIP_TRY(hipMemset(ctx->d_zero64, 0, sizeof(int64_t)));
    // Allow the order-bucketed kernel the full LDS of a CU.
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_fpm_cdet),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kMaxLds)));
    refresh_scalars(ctx);
    *out_ctx = ctx;
    return 0;
}

int cdet_ctx_destroy(cdet_ctx* ctx) {
    if (!ctx) return 0;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& t : ctx->tab) {
        if (t.alias < 0) {
            (void)hipFree(t.upper);
            (void)hipFree(t.lower);
            (void)hipFree(t.ucorner);
            (void)hipFree(t.lcorner);
            for (auto& b : t.split)
                for (float* f : b) (void)hipFree(f);
        }
    }
    (void)hipFree(ctx->d_masks);
    (void)hipFree(ctx->d_popoff);
    (void)hipFree(ctx->d_fallbacks);
    (void)hipFree(ctx->d_verts);
    (void)hipFree(ctx->d_gup);
    (void)hipFree(ctx->d_gdn);
    (void)hipFree(ctx->d_out);
    (void)hipFree(ctx->d_status);
    (void)hipFree(ctx->d_one);
    (void)hipFree(ctx->d_zero64);
    (void)hipFree(ctx->d_gscratch);
    if (ctx->host_plan) cdet_batch_plan_destroy(ctx->host_plan);
    (void)hipFree(ctx->d_hv);
    (void)hipFree(ctx->d_hout);
    (void)hipFree(ctx->d_hstatus);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

void* cdet_ctx_stream(cdet_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

static int table_sizes(const cdet_ctx* ctx, size_t& ncoef, size_t& ncorner) {
    const size_t S = static_cast<size_t>(ctx->p.N1) * ctx->p.N2;
    ncorner = 4 * S * S;
    ncoef = ncorner * static_cast<size_t>(ctx->p.cheb_nu) * ctx->p.cheb_nv;
    return 0;
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
    for (auto& b : t.split)
        for (float*& f : b) {
            (void)hipFree(f);
            f = nullptr;
        }
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
        for (auto& b : t.split)
            for (float* f : b) (void)hipFree(f);
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

static int plain_eval(cdet_ctx* ctx, const cdet_vertex* v, int n, int mode, double* out, size_t out_count) {
    if (!tables_ready(ctx)) return fail(CDET_ERR_TABLES, "Resident Green-function tables have not been uploaded");
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(ctx->d_verts, v, sizeof(double) * 4 * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_pairs_plain, dim3(2 * n * n), dim3(64), 0, s, ctx->ks, ctx->d_verts, n, ctx->d_gup,
                       ctx->d_gdn);
    HIP_TRY(hipGetLastError());
    int rc = launch_fpm(ctx, n, mode, ctx->d_one, 1, ctx->d_zero64, ctx->d_zero64, ctx->d_gup, ctx->d_gdn, ctx->d_out,
                        ctx->d_status, s);
    if (rc) return rc;
    int32_t st = 0;
    HIP_TRY(hipMemcpyAsync(out, ctx->d_out, sizeof(double) * out_count, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&st, ctx->d_status, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    ctx->calls += 1;
    if (st & 1) {
        if (mode == 1) return fail(CDET_ERR_NONFINITE, "Error (integrand): Value is not finite");
        return fail(CDET_ERR_NONFINITE, "Error (all-subset integrand): Value is not finite");
    }
    return 0;
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

int cdet_get_counters(cdet_ctx* ctx, cdet_counters* out) {
    if (!ctx || !out) return fail(CDET_ERR_ARGUMENT, "null argument");
    uint64_t fb = 0;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemcpy(&fb, ctx->d_fallbacks, sizeof(uint64_t), hipMemcpyDeviceToHost));
    out->integrand_calls = ctx->calls;
    out->fpm_fallbacks = fb;
    out->green_pair_cache_hits = 0;
    out->green_pair_cache_misses = 0;
    return 0;
}

int cdet_reset_counters(cdet_ctx* ctx) {
    if (!ctx) return fail(CDET_ERR_ARGUMENT, "null argument");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemset(ctx->d_fallbacks, 0, sizeof(uint64_t)));
    ctx->calls = 0;
    return 0;
}

}  // extern "C"

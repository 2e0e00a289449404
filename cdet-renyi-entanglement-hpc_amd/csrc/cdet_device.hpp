// cdet_device.hpp -- gfx950 device building blocks of the CDet integrand path.
//
// Everything here is FP64 and follows the reference's arithmetic operation by
// operation (compiled with -ffp-contract=off; fma only where the reference
// calls std::fma), so a value computed here is bit-identical to the
// reference's CPU value for the same inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cdet {

constexpr int kMaxOrder = 14;
constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Double-double (include/cdet/utility/types.h:30-150): Dekker/Knuth
// error-free transforms in the reference's sequence.
// ---------------------------------------------------------------------------
struct DD {
    double hi;
    double lo;
};
__device__ __host__ inline double dd_quick_two_sum(double a, double b, double& err) {
    const double s = a + b;
    err = b - (s - a);
    return s;
}
__device__ __host__ inline double dd_two_sum(double a, double b, double& err) {
    const double s = a + b;
    const double bb = s - a;
    err = (a - (s - bb)) + (b - bb);
    return s;
}
__device__ inline DD dd_sub(DD x, DD y) {
    double s2, t2;
    double s1 = dd_two_sum(x.hi, -y.hi, s2);
    const double t1 = dd_two_sum(x.lo, -y.lo, t2);
    s2 += t1;
    s1 = dd_quick_two_sum(s1, s2, s2);
    s2 += t2;
    DD r;
    r.hi = dd_quick_two_sum(s1, s2, r.lo);
    return r;
}
__device__ inline DD dd_mul(DD x, DD y) {
    const double p = x.hi * y.hi;
    double e = __builtin_fma(x.hi, y.hi, -p);
    e += x.hi * y.lo + x.lo * y.hi + x.lo * y.lo;
    DD r;
    r.hi = dd_quick_two_sum(p, e, r.lo);
    return r;
}

// ---------------------------------------------------------------------------
// Vertex decode and lookup descriptor (include/cdet/g_funcs.h:300-342,
// :609-653; include/cdet/integrand_replica_randomly_sampled.h:142-153, :652-662).
// ---------------------------------------------------------------------------
struct Vtx {
    double tau;
    double tau_unit;
    int x;
    int y;
    int replica;
};

__device__ inline int coord_to_index(double coordinate, int extent) {
    // floor of a double is exact, as is the reference's long-double floor.
    int index = static_cast<int>(floor(coordinate));
    if (index < 0) index = 0;
    if (index >= extent) index = extent - 1;
    return index;
}

__device__ inline Vtx decode_vertex(const double* v4, double beta, int N1, int N2, int rc) {
    Vtx s;
    s.tau_unit = v4[0];
    s.tau = beta * s.tau_unit;
    s.x = coord_to_index(static_cast<double>(N1) * v4[1], N1);
    s.y = coord_to_index(static_cast<double>(N2) * v4[2], N2);
    if (rc <= 1) {
        s.replica = 0;
    } else {
        int r = coord_to_index(static_cast<double>(rc) * v4[3], rc);
        s.replica = r < 0 ? 0 : (r > rc - 1 ? rc - 1 : r);
    }
    return s;
}

__device__ inline double clamp_unit(double v) {
    if (v < -1.0) return -1.0;
    if (v > 1.0) return 1.0;
    return v;
}

__device__ __host__ inline int64_t sitepair_index(int ra, int rb, int x1, int y1, int x2, int y2, int N1, int N2) {
    const int64_t r = static_cast<int64_t>(ra) * 2 + rb;
    return ((((r * N1 + x1) * N2 + y1) * N1 + x2) * N2) + y2;
}

struct Desc {
    int upper;
    int corner;
    int64_t sitepair;
    double u;
    double v;
};

__device__ inline Desc make_desc(const Vtx& vi, const Vtx& vj, bool upper, double beta, double branch_eps, int N1,
                                 int N2) {
    Desc d;
    d.upper = upper ? 1 : 0;
    d.corner = 0;
    d.u = 0.0;
    d.v = 0.0;
    d.sitepair = sitepair_index(vi.replica, vj.replica, vi.x, vi.y, vj.x, vj.y, N1, N2);
    const double t1 = vi.tau;
    const double t2 = vj.tau;
    if (upper) {
        if (t1 < branch_eps) {
            d.corner = 1;
            return d;
        }
        d.u = clamp_unit(2.0 * t1 / beta - 1.0);
        d.v = clamp_unit(2.0 * t2 / t1 - 1.0);
    } else {
        if (t2 < branch_eps) {
            d.corner = 1;
            return d;
        }
        d.u = clamp_unit(2.0 * t2 / beta - 1.0);
        d.v = clamp_unit(2.0 * t1 / t2 - 1.0);
    }
    return d;
}

// ---------------------------------------------------------------------------
// Clenshaw over one contiguous coefficient row (include/cdet/g_funcs.h:426-463).
// ---------------------------------------------------------------------------
__device__ inline double clenshaw_row(const double* c, int degree, double x) {
    if (degree <= 0) return degree < 0 ? 0.0 : c[0];
    if (degree == 1) return c[0] + c[1] * x;
    const double two_x = 2.0 * x;
    double b1 = 0.0, b2 = 0.0;
    for (int k = degree; k >= 1; --k) {
        const double bk = c[k] + two_x * b1 - b2;
        b2 = b1;
        b1 = bk;
    }
    return c[0] + b1 * x - b2;
}

// Incremental Clenshaw over u that consumes row values r[nu-1], ..., r[0] in
// that order (the order in which the reference's final Clenshaw reads them,
// include/cdet/g_funcs.h:748 -> :450-461).
struct UClenshaw {
    double b1, b2, r1, two_u, u;
    int degree;
    __device__ inline void init(double uu, int nu) {
        b1 = 0.0;
        b2 = 0.0;
        r1 = 0.0;
        u = uu;
        two_u = 2.0 * uu;
        degree = nu - 1;
    }
    // Feed row value of index a (a descending); returns the final value at a == 0.
    __device__ inline double feed(int a, double r) {
        if (degree >= 2) {
            if (a >= 1) {
                const double bk = r + two_u * b1 - b2;
                b2 = b1;
                b1 = bk;
                return 0.0;
            }
            return r + b1 * u - b2;
        }
        if (degree == 1) {
            if (a == 1) {
                r1 = r;
                return 0.0;
            }
            return r + r1 * u;
        }
        return r;
    }
};

// ---------------------------------------------------------------------------
// Principal minors: the pivoted-determinant fallback and the
// workgroup-cooperative fast principal minors
// (include/cdet/fast_principle_minors.h:39-246).  All intermediates are
// formed with the reference's operations; only the iteration order across
// independent elements differs, which cannot change any value.
// ---------------------------------------------------------------------------
// Pivoted determinant of A[mask, mask] with the row elimination of
// include/cdet/fast_principle_minors.h:39-115, the k x k working matrix
// spread over a group of 16 lanes (lanes gbase .. gbase + 15 of the wave):
// lane r holds row r in NM
// registers, so a wave runs four masks at once at a few dozen registers per
// lane instead of NM^2 doubles per mask on one lane.  Every lane of the wave
// must call it (the exchanges are wave-wide); groups whose mask is 0 return 1.
//   pivot: the reference scans rows col+1.. and takes a row only if its |a|
//     is strictly greater than the best so far, i.e. the first row of the
//     largest |a| among those greater than the pivot candidate's (NaN rows
//     never qualify; a NaN candidate keeps its row).  Each lane offers |a| if
//     it qualifies, else -1, and a butterfly keeps the larger value, the
//     lower row on ties;
//   swap: rows col and prow exchange their entries from column col on;
//   elimination: each lower row forms f = a_rc / pivot and, if f != 0,
//     a_rc' = fma(-f, pivot row c', a_rc') for the columns to its right,
//     with the pivot row broadcast from its lane.
// The arithmetic and the early zero are the reference's, value for value;
// the determinant (det *= pivot, sign flips) is formed in every lane.
template <int NM>
__device__ inline double det_minor_rows(const double* A, int n, uint32_t mask, int r, int gbase) {
    static_assert(NM <= 16, "one row per lane of a 16-lane group");
    const int k = __builtin_popcount(mask);
    // row index of this lane and the column indices (the set bits of mask)
    int cidx[NM];
    {
        uint32_t mm = mask;
#pragma unroll
        for (int c = 0; c < NM; ++c) {
            cidx[c] = mm ? __builtin_ctz(mm) : 0;
            mm &= mm - 1u;
        }
    }
    int ridx = 0;
#pragma unroll
    for (int c = 0; c < NM; ++c)
        if (c == r) ridx = cidx[c];
    const bool row_live = r < k;
    double W[NM];
#pragma unroll
    for (int c = 0; c < NM; ++c) W[c] = (row_live && c < k) ? A[ridx * n + cidx[c]] : 0.0;
    double det = 1.0;
    int sign = 1;
    bool zero = false;
#pragma unroll
    for (int col = 0; col < NM; ++col) {
        const bool act = col < k && !zero;  // uniform within the group
        // pivot candidate |a_col,col| from lane col, then the qualifying rows
        const double vcol = __shfl(fabs(W[col]), gbase + col);
        const double v = fabs(W[col]);
        double best = (act && r > col && r < k && v > vcol) ? v : -1.0;
        int brow = r;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const double ob = __shfl_xor(best, o);
            const int orow = __shfl_xor(brow, o);
            if (ob > best || (ob == best && orow < brow)) {
                best = ob;
                brow = orow;
            }
        }
        const int prow = best >= 0.0 ? brow : col;
        const double pabs = best >= 0.0 ? best : vcol;
        if (act && pabs == 0.0) zero = true;
        const bool go = act && !zero;
        // rows col and prow exchange their entries from column col on
        const bool swap = go && prow != col;
        const int partner = (r == col) ? prow : (r == prow ? col : r);
#pragma unroll
        for (int c = col; c < NM; ++c) {
            const double o = __shfl(W[c], gbase + partner);
            if (swap) W[c] = o;
        }
        if (swap) sign = -sign;
        // the pivot row, broadcast from lane col
        double P[NM];
#pragma unroll
        for (int c = col; c < NM; ++c) P[c] = __shfl(W[c], gbase + col);
        if (go) {
            const double piv = P[col];
            det *= piv;
            if (r > col && r < k) {
                const double f = W[col] / piv;
                if (f != 0.0) {
                    W[col] = 0.0;
#pragma unroll
                    for (int c = col + 1; c < NM; ++c)
                        if (c < k) W[c] = __builtin_fma(-f, P[c], W[c]);
                }
            }
        }
    }
    if (k == 0) return 1.0;
    if (zero) return 0.0;
    return sign < 0 ? -det : det;
}

// Maximum over the workgroup: butterfly within each wave, then across the
// waves through red[] (exact for any order).
__device__ inline double block_max(double v, double* red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    const int nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < nw; ++w) r = red[w] > r ? red[w] : r;
    __syncthreads();
    return r;
}

// x / d for 0 <= x < 2^22 and 1 <= d < 2^10, with inv_d = 1.0f / d: the float
// quotient is within one of the true one and is corrected exactly.
__device__ inline int div_small(int x, int d, float inv_d) {
    int q = static_cast<int>(static_cast<float>(x) * inv_d);
    if (q * d > x) --q;
    else if ((q + 1) * d <= x) ++q;
    return q;
}

// out[mask] = det(A[mask, mask]) for every mask (out[0] = 1).  With
// brute = false this is the Schur-complement tree of A + pI and returns 1,
// leaving out undefined, as soon as a pivot |A00 + p| falls to
// tol * max|A| (the reference then switches to pivoted determinants,
// include/cdet/fast_principle_minors.h:180-186); with brute = true the
// pivoted determinants, formed beforehand by k_fallback_dets, are taken
// from pre[mask].  Scratch: bufA/bufB hold fpm_list_size(n) >= 2^n doubles,
// inv 2^(n-2), red blockDim.
template <bool BRUTE>
__device__ inline int fpm_block(const double* A, int n, double tol, double thresh, double* out, double* bufA,
                                double* bufB, double* inv, double* red, int* flag, const double* pre = nullptr) {
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const uint32_t sc = 1u << n;
    const int msize = n * n;

    double lmax = 0.0;
    for (int i = tid; i < msize; i += nt) {
        const double e = fabs(A[i]);
        if (e > lmax) lmax = e;
    }
    const double max_abs = block_max(lmax, red);
    const bool explicit_shift = !(thresh < 0.0);
    const double p = explicit_shift ? thresh : max_abs / static_cast<double>(msize);
    if (max_abs == 0.0 && p == 0.0) {
        for (uint32_t m = tid; m < sc; m += nt) out[m] = (m == 0) ? 1.0 : 0.0;
        __syncthreads();
        return 0;
    }
    if (BRUTE) {
        for (uint32_t m = tid; m < sc; m += nt) out[m] = (m == 0) ? 1.0 : pre[m];
        __syncthreads();
        return 0;
    }
    const double fb_eps = (tol > 0.0 && max_abs > 0.0) ? tol * max_abs : 0.0;

    // shifted_minors: 1, p + A00, p, p, ... (:160-163)
    for (int i = tid; i < msize; i += nt) bufA[i] = A[i];
    for (uint32_t m = tid; m < sc; m += nt) out[m] = (m == 0) ? 1.0 : (m == 1 ? p + A[0] : p);
    if (tid == 0) *flag = 0;
    __syncthreads();

    double* oldl = bufA;
    double* newl = bufB;
    for (int l = 0; l < n - 1; ++l) {
        const int os = n - l;
        const int oms = os * os;
        const int ns = os - 1;
        const int nms = ns * ns;
        const int sets = 1 << l;
        for (int s = tid; s < sets; s += nt) {
            const double a00 = oldl[s * oms];
            // minors of the matrices generated one level up (:210-213)
            if (l >= 1) out[(1u << l) + s] += a00;
            const double pd = a00 + p;
            const double pda = fabs(pd);
            if (pda == 0.0 || (fb_eps > 0.0 && pda <= fb_eps)) *flag = 1;
            inv[s] = 1.0 / pd;
        }
        __syncthreads();
        if (*flag) return 1;
        const int items = sets * nms;
        const float inv_nms = 1.0f / static_cast<float>(nms);
        const float inv_ns = 1.0f / static_cast<float>(ns);
        // item it = s nms + rem, rem = r ns + c; from the thread's first item
        // the indices step by nt as (ds, drem = dr ns + dc) with carries
        const int ds = div_small(nt, nms, inv_nms);
        const int drem = nt - ds * nms;
        const int dr = div_small(drem, ns, inv_ns);
        const int dc = drem - dr * ns;
        int s = div_small(tid, nms, inv_nms);
        int rem = tid - s * nms;
        int r = div_small(rem, ns, inv_ns);
        int c = rem - r * ns;
        for (int it = tid; it < items; it += nt) {
            const int src = s * oms;
            const double f = oldl[src + (r + 1) * os] * inv[s];
            const double base = oldl[src + (r + 1) * os + (c + 1)];
            newl[s * nms + rem] = base;
            newl[(s + sets) * nms + rem] = __builtin_fma(-oldl[src + 1 + c], f, base);
            s += ds;
            rem += drem;
            r += dr;
            c += dc;
            if (c >= ns) {
                c -= ns;
                ++r;
            }
            if (rem >= nms) {
                rem -= nms;
                r -= ns;
                ++s;
            }
        }
        __syncthreads();
        double* t = oldl;
        oldl = newl;
        newl = t;
    }
    if (n >= 2) {
        const uint32_t top = 1u << (n - 1);
        for (uint32_t t = tid; t < top; t += nt) out[top + t] += oldl[t];
    }
    __syncthreads();
    // Prefix products (:217-223): final[m] = r[m] * final[m - topbit(m)], the
    // running product over the bits of m from the lowest upward (final of a
    // one-bit mask = r[m]).  Thread t < T = min(2^n, nt) forms final[t] as
    // that running product, then the masks h * T + t in increasing h, each one
    // product with final[(h - topbit(h)) * T + t], which the same thread wrote
    // before (nt is a power of two).
    {
        const uint32_t T = sc < static_cast<uint32_t>(nt) ? sc : static_cast<uint32_t>(nt);
        if (static_cast<uint32_t>(tid) < T) {
            const uint32_t t = static_cast<uint32_t>(tid);
            double val = 1.0;
            if (t != 0) {
                uint32_t pre = t & (~t + 1u);
                uint32_t rest = t ^ pre;
                val = out[pre];
                while (rest) {
                    const uint32_t b = rest & (~rest + 1u);
                    rest ^= b;
                    pre |= b;
                    val = out[pre] * val;
                }
            }
            bufA[t] = val;
            const uint32_t hn = sc / T;
            for (uint32_t h = 1; h < hn; ++h) {
                const uint32_t m = h * T + t;
                const uint32_t r = (h ^ (1u << (31 - __builtin_clz(h)))) * T + t;
                bufA[m] = r == 0 ? out[m] : out[m] * bufA[r];
            }
        }
    }
    __syncthreads();
    // Removal of the uniform shift, highest bit first (:225-239): level l
    // takes bufA[m] = fma(-p, bufA[m - 2^(l-1)], bufA[m]) for the masks with
    // bit l-1 set.  Two levels per pass: a thread holds the four masks that
    // differ in bits l-1 and l-2 and applies level l, then level l-1, in
    // registers.  The last pass writes the minors to out.
    for (int l = n; l > 0;) {
        const bool last = l <= 2;
        double* dst = last ? out : bufA;
        if (l >= 2) {
            const uint32_t H = 1u << (l - 1);
            const uint32_t L = 1u << (l - 2);
            for (uint32_t g = tid; g < (sc >> 2); g += nt) {
                const uint32_t x = ((g >> (l - 2)) << l) | (g & (L - 1));
                const double b0 = bufA[x];
                double bl = bufA[x + L];
                double bh = bufA[x + H];
                double bhl = bufA[x + H + L];
                bh = __builtin_fma(-p, b0, bh);
                bhl = __builtin_fma(-p, bl, bhl);
                bl = __builtin_fma(-p, b0, bl);
                bhl = __builtin_fma(-p, bh, bhl);
                dst[x] = b0;
                dst[x + L] = bl;
                dst[x + H] = bh;
                dst[x + H + L] = bhl;
            }
            l -= 2;
        } else {
            for (uint32_t q = tid; q < (sc >> 1); q += nt) {
                const double b0 = bufA[2 * q];
                dst[2 * q] = b0;
                dst[2 * q + 1] = __builtin_fma(-p, b0, bufA[2 * q + 1]);
            }
            l -= 1;
        }
        __syncthreads();
    }
    return 0;
}

// Size (doubles) of each FPM list buffer for order n.
__host__ __device__ inline int fpm_list_size(int n) {
    int best = n * n;
    for (int l = 0; l < n - 1; ++l) {
        const int ns = n - l - 1;
        const int c = (1 << (l + 1)) * ns * ns;
        if (c > best) best = c;
    }
    return best;
}

}  // namespace cdet

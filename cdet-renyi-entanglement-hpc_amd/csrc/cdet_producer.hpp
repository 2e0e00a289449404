// cdet_producer.hpp -- FP64 replica-Green table construction on the GPU
// (included by cdet_kernels.hip).
//
// Reference: load_data_into_resident_tables (include/cdet/load_data.h:49-99),
// build_green_function_table / build_ordinary_replica00_green_function_table
// (include/cdet/replica_chebyshev_tables.h:592-793) and the stable graded
// kernel (include/cdet/replica_free_greens.h:423-534).
//
// Host: lattice Hamiltonian, cyclic Jacobi eigendecomposition (S x S, once)
// and the sector-mixing matrix S_0 = V^T Omega_0^{-1} V.
// Device, per Chebyshev node (one workgroup, matrices in LDS):
//   middle = diag(e^{-theta c}) S_0 diag(e^{-(1-theta) c}) + diag(e^{x-c})
//   (conjugate transpose of S_0 on the lower branch), solved against the
//   diagonal right-hand side by pivoted complex Gauss-Jordan, scaled by the
//   left factor, and rotated to sites G = V K V^T.
// The second replica sector is the complex conjugate of the first
// (S_1 = conj(S_0) with everything else real), so one solve per node gives
// both: (g0+g1)/2 = Re g0, i(g0-g1)/2 = -Im g0, -i(g0-g1)/2 = Im g0.
// Samples land at [sitepair][node] in the table itself and are transformed in
// place, per sitepair, by the separable cosine transform
// (include/cdet/replica_chebyshev_tables.h:552-590).

namespace {

constexpr int kProdMaxSites = 64;  // LDS-resident solve: 2 * 64^2 complex

struct ProdArgs {
    int S, N1, N2, nu, nv, ordinary, upper;
    double beta;
    const double* evals;
    const double* V;       // V[site * S + p]
    const double* mix_re;  // S_0, row-major
    const double* mix_im;
    const double* node_u;
    const double* node_v;
    double* dst;        // samples [sitepair][node]
    double* corner;     // corner values [sitepair] (corner launch)
    int32_t* err;       // 1 = singular pivot
};

__device__ inline void cmul_(double ar, double ai, double br, double bi, double& r, double& i) {
    r = ar * br - ai * bi;
    i = ar * bi + ai * br;
}

__device__ inline void cdiv_(double ar, double ai, double br, double bi, double& r, double& i) {
    const double den = br * br + bi * bi;
    r = (ar * br + ai * bi) / den;
    i = (ai * br - ar * bi) / den;
}

__global__ __launch_bounds__(256) void k_node_green(ProdArgs a, int node_begin, int corner_mode) {
    extern __shared__ double sh[];
    const int S = a.S;
    const int S2 = S * S;
    double* Mr = sh;
    double* Mi = Mr + S2;
    double* Rr = Mi + S2;
    double* Ri = Rr + S2;
    double* left = Ri + S2;
    double* colr = left + S;
    double* coli = colr + S;
    __shared__ int s_piv;
    __shared__ double s_pabs[256];
    __shared__ int s_prow[256];
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const bool upper = a.upper != 0;
    const int node = node_begin + blockIdx.x;
    double tau = 0.0, taup = 0.0;
    if (!corner_mode) {
        const int i = node / a.nv;
        const int j = node - i * a.nv;
        const double cu = a.node_u[i];
        const double cv = a.node_v[j];
        if (upper) {
            tau = a.beta * (1.0 + cu) / 2.0;
            taup = tau * (1.0 + cv) / 2.0;
        } else {
            taup = a.beta * (1.0 + cu) / 2.0;
            tau = taup * (1.0 + cv) / 2.0;
        }
    }
    const double tf = tau / a.beta;
    const double tpf = taup / a.beta;
    const double theta = upper ? (tf + tpf) / 2.0 : 1.0 - (tf + tpf) / 2.0;
    // middle and diagonal right-hand side (replica_free_greens.h:445-476)
    for (int e = tid; e < S2; e += nt) {
        const int r = e / S;
        const int c = e - r * S;
        const double xr = a.beta * a.evals[r];
        const double xc = a.beta * a.evals[c];
        const double sxr = upper ? xr : -xr;
        const double sxc = upper ? xc : -xc;
        const double cr = fmax(0.0, sxr);
        const double cc = fmax(0.0, sxc);
        const double rs = exp(-theta * cr);
        const double cs = exp(-(1.0 - theta) * cc);
        double mr, mi;
        if (upper) {
            mr = a.mix_re[r * S + c];
            mi = a.mix_im[r * S + c];
        } else {
            mr = a.mix_re[c * S + r];
            mi = -a.mix_im[c * S + r];
        }
        Mr[e] = (rs * mr) * cs;
        Mi[e] = (rs * mi) * cs;
        if (r == c) Mr[e] += exp(sxr - cr);
        Rr[e] = 0.0;
        Ri[e] = 0.0;
    }
    // the diagonal of R is set below by other threads than the ones that
    // cleared it above
    __syncthreads();
    for (int r = tid; r < S; r += nt) {
        const double x = a.beta * a.evals[r];
        const double sx = upper ? x : -x;
        const double c = fmax(0.0, sx);
        double le, re;
        if (upper) {
            le = (1.0 - tf) * x - (1.0 - theta) * c;
            re = tpf * x - theta * c;
        } else {
            le = tf * sx - (1.0 - theta) * c;
            re = (1.0 - tpf) * sx - theta * c;
        }
        left[r] = upper ? exp(le) : -exp(le);
        Rr[r * S + r] = exp(re);
    }
    __syncthreads();
    // Pivoted complex Gauss-Jordan (fermionic_graded_complex.h:304-361).
    for (int col = 0; col < S; ++col) {
        double best = -1.0;
        int brow = S;
        for (int r = col + tid; r < S; r += nt) {
            const double v = sqrt(Mr[r * S + col] * Mr[r * S + col] + Mi[r * S + col] * Mi[r * S + col]);
            if (v > best) {
                best = v;
                brow = r;
            }
        }
        s_pabs[tid] = best;
        s_prow[tid] = brow;
        __syncthreads();
        if (tid == 0) {
            double pb = -1.0;
            int pr = S;
            for (int t = 0; t < nt; ++t) {
                if (s_pabs[t] > pb || (s_pabs[t] == pb && s_prow[t] < pr)) {
                    pb = s_pabs[t];
                    pr = s_prow[t];
                }
            }
            if (pb <= 1e-12) atomicOr(a.err, 1);
            s_piv = pr;
        }
        __syncthreads();
        const int prow = s_piv;
        if (prow != col) {
            for (int j = tid; j < S; j += nt) {
                double t;
                t = Mr[col * S + j]; Mr[col * S + j] = Mr[prow * S + j]; Mr[prow * S + j] = t;
                t = Mi[col * S + j]; Mi[col * S + j] = Mi[prow * S + j]; Mi[prow * S + j] = t;
                t = Rr[col * S + j]; Rr[col * S + j] = Rr[prow * S + j]; Rr[prow * S + j] = t;
                t = Ri[col * S + j]; Ri[col * S + j] = Ri[prow * S + j]; Ri[prow * S + j] = t;
            }
            __syncthreads();
        }
        const double pr_ = Mr[col * S + col];
        const double pi_ = Mi[col * S + col];
        __syncthreads();
        for (int j = tid; j < 2 * S; j += nt) {
            double* xr = j < S ? Mr : Rr;
            double* xi = j < S ? Mi : Ri;
            const int jj = j < S ? j : j - S;
            if (j < S && jj < col) continue;
            double r, i;
            cdiv_(xr[col * S + jj], xi[col * S + jj], pr_, pi_, r, i);
            xr[col * S + jj] = r;
            xi[col * S + jj] = i;
        }
        for (int r = tid; r < S; r += nt) {
            colr[r] = Mr[r * S + col];
            coli[r] = Mi[r * S + col];
        }
        __syncthreads();
        for (int e = tid; e < S * 2 * S; e += nt) {
            const int r = e / (2 * S);
            const int j = e - r * 2 * S;
            if (r == col) continue;
            const double fr = colr[r], fi = coli[r];
            if (fr == 0.0 && fi == 0.0) continue;
            double* xr = j < S ? Mr : Rr;
            double* xi = j < S ? Mi : Ri;
            const int jj = j < S ? j : j - S;
            if (j < S && jj < col) continue;
            double pr2, pi2;
            cmul_(fr, fi, xr[col * S + jj], xi[col * S + jj], pr2, pi2);
            xr[r * S + jj] -= pr2;
            xi[r * S + jj] -= pi2;
        }
        __syncthreads();
    }
    // K = diag(left) X, then T = V K (into M), G = T V^T (into R).
    for (int e = tid; e < S2; e += nt) {
        const int r = e / S;
        Rr[e] *= left[r];
        Ri[e] *= left[r];
    }
    __syncthreads();
    for (int e = tid; e < S2; e += nt) {
        const int i = e / S;
        const int q = e - i * S;
        double tr = 0.0, ti = 0.0;
        for (int p = 0; p < S; ++p) {
            const double v = a.V[i * S + p];
            tr = __builtin_fma(v, Rr[p * S + q], tr);
            ti = __builtin_fma(v, Ri[p * S + q], ti);
        }
        Mr[e] = tr;
        Mi[e] = ti;
    }
    __syncthreads();
    for (int e = tid; e < S2; e += nt) {
        const int i = e / S;
        const int jx = e - i * S;
        double gr = 0.0, gi = 0.0;
        for (int q = 0; q < S; ++q) {
            const double v = a.V[jx * S + q];
            gr = __builtin_fma(Mr[i * S + q], v, gr);
            gi = __builtin_fma(Mi[i * S + q], v, gi);
        }
        Rr[e] = gr;
        Ri[e] = gi;
    }
    __syncthreads();
    const int64_t stride = corner_mode ? 1 : static_cast<int64_t>(a.nu) * a.nv;
    double* out = corner_mode ? a.corner : a.dst + node;
    if (a.ordinary) {
        for (int e = tid; e < S2; e += nt) out[static_cast<int64_t>(e) * stride] = Rr[e];
        return;
    }
    const int SP = 4 * S2;
    for (int f = tid; f < SP; f += nt) {
        int g = f;
        const int y2 = g % a.N2;
        g /= a.N2;
        const int x2 = g % a.N1;
        g /= a.N1;
        const int y1 = g % a.N2;
        g /= a.N2;
        const int x1 = g % a.N1;
        g /= a.N1;
        const int rb = g & 1;
        const int ra = g >> 1;
        const int s1 = y1 * a.N1 + x1;
        const int s2 = y2 * a.N1 + x2;
        const double re = Rr[s1 * S + s2];
        const double im = Ri[s1 * S + s2];
        const double v = (ra == rb) ? re : (ra == 0 ? -im : im);
        out[static_cast<int64_t>(f) * stride] = v;
    }
}

// Lattices of 65..100 sites: the S x S complex matrix alone fills the LDS, so
// the solve is an in-place Gauss-Jordan inversion with partial pivoting
// (row swaps recorded, columns unscrambled in reverse at the end), and the two
// basis rotations stage their S x S products through a per-workgroup slot of
// global scratch.  One workgroup per CU; each loops over nodes.
//   K = diag(left) M^{-1} diag(e),  T = V K  (to scratch, then back to LDS),
//   G = T V^T (to scratch), then the replica entries as in k_node_green.
constexpr int kProdBigMaxSites = 100;
constexpr int kProdBigThreads = 1024;

__device__ inline void big_pivot_reduce(double& best, int& brow, double* red_v, int* red_i) {
    // wave maximum, larger value first, then smaller row
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(best, off);
        const int orow = __shfl_xor(brow, off);
        if (ov > best || (ov == best && orow < brow)) {
            best = ov;
            brow = orow;
        }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red_v[w] = best;
        red_i[w] = brow;
    }
}

__global__ __launch_bounds__(kProdBigThreads) void k_node_green_big(ProdArgs a, int node_begin, int node_end,
                                                                    int corner_mode, double* __restrict__ tscr) {
    extern __shared__ double sh[];
    const int S = a.S;
    const int S2 = S * S;
    double* Ar = sh;
    double* Ai = Ar + S2;
    double* colr = Ai + S2;  // pivot column snapshot; later left[] and e[]
    double* coli = colr + S;
    double* red_v = coli + S;                                  // one per wave
    int* red_i = reinterpret_cast<int*>(red_v + kProdBigThreads / 64);
    int* perm = red_i + kProdBigThreads / 64;
    int* s_piv = perm + S;
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const int nw = nt >> 6;
    const bool upper = a.upper != 0;
    double* tr_g = tscr + static_cast<int64_t>(blockIdx.x) * 2 * S2;
    double* ti_g = tr_g + S2;
    const int last = corner_mode ? 1 : node_end;
    for (int node = corner_mode ? 0 : node_begin + blockIdx.x; node < last; node += gridDim.x) {
        double tau = 0.0, taup = 0.0;
        if (!corner_mode) {
            const int i = node / a.nv;
            const int j = node - i * a.nv;
            const double cu = a.node_u[i];
            const double cv = a.node_v[j];
            if (upper) {
                tau = a.beta * (1.0 + cu) / 2.0;
                taup = tau * (1.0 + cv) / 2.0;
            } else {
                taup = a.beta * (1.0 + cu) / 2.0;
                tau = taup * (1.0 + cv) / 2.0;
            }
        }
        const double tf = tau / a.beta;
        const double tpf = taup / a.beta;
        const double theta = upper ? (tf + tpf) / 2.0 : 1.0 - (tf + tpf) / 2.0;
        // middle (replica_free_greens.h:445-476)
        for (int e = tid; e < S2; e += nt) {
            const int r = e / S;
            const int c = e - r * S;
            const double xr = a.beta * a.evals[r];
            const double xc = a.beta * a.evals[c];
            const double sxr = upper ? xr : -xr;
            const double sxc = upper ? xc : -xc;
            const double cr = fmax(0.0, sxr);
            const double cc = fmax(0.0, sxc);
            const double rs = exp(-theta * cr);
            const double cs = exp(-(1.0 - theta) * cc);
            double mr, mi;
            if (upper) {
                mr = a.mix_re[r * S + c];
                mi = a.mix_im[r * S + c];
            } else {
                mr = a.mix_re[c * S + r];
                mi = -a.mix_im[c * S + r];
            }
            Ar[e] = (rs * mr) * cs;
            Ai[e] = (rs * mi) * cs;
            if (r == c) Ar[e] += exp(sxr - cr);
        }
        __syncthreads();
        // in-place inversion
        for (int col = 0; col < S; ++col) {
            double best = -1.0;
            int brow = S;
            for (int r = col + tid; r < S; r += nt) {
                const double v = sqrt(Ar[r * S + col] * Ar[r * S + col] + Ai[r * S + col] * Ai[r * S + col]);
                if (v > best) {
                    best = v;
                    brow = r;
                }
            }
            big_pivot_reduce(best, brow, red_v, red_i);
            __syncthreads();
            if (tid == 0) {
                double pb = red_v[0];
                int pr = red_i[0];
                for (int w = 1; w < nw; ++w)
                    if (red_v[w] > pb || (red_v[w] == pb && red_i[w] < pr)) {
                        pb = red_v[w];
                        pr = red_i[w];
                    }
                if (pb <= 1e-12 || pr >= S) {
                    atomicOr(a.err, 1);
                    pr = col;
                }
                *s_piv = pr;
                perm[col] = pr;
            }
            __syncthreads();
            const int prow = *s_piv;
            if (prow != col) {
                for (int j = tid; j < S; j += nt) {
                    double t;
                    t = Ar[col * S + j]; Ar[col * S + j] = Ar[prow * S + j]; Ar[prow * S + j] = t;
                    t = Ai[col * S + j]; Ai[col * S + j] = Ai[prow * S + j]; Ai[prow * S + j] = t;
                }
                __syncthreads();
            }
            const double pr_ = Ar[col * S + col];
            const double pi_ = Ai[col * S + col];
            __syncthreads();
            // pivot row / pivot, with the pivot entry replaced by 1 first
            for (int j = tid; j < S; j += nt) {
                const double xr = j == col ? 1.0 : Ar[col * S + j];
                const double xi = j == col ? 0.0 : Ai[col * S + j];
                double r, i;
                cdiv_(xr, xi, pr_, pi_, r, i);
                Ar[col * S + j] = r;
                Ai[col * S + j] = i;
            }
            for (int r = tid; r < S; r += nt) {
                colr[r] = Ar[r * S + col];
                coli[r] = Ai[r * S + col];
            }
            __syncthreads();
            for (int e = tid; e < S2; e += nt) {
                const int r = e / S;
                if (r == col) continue;
                const int j = e - r * S;
                const double fr = colr[r], fi = coli[r];
                double pr2, pi2;
                cmul_(fr, fi, Ar[col * S + j], Ai[col * S + j], pr2, pi2);
                Ar[e] = (j == col ? 0.0 : Ar[e]) - pr2;
                Ai[e] = (j == col ? 0.0 : Ai[e]) - pi2;
            }
            __syncthreads();
        }
        for (int k = S - 1; k >= 0; --k) {
            const int q = perm[k];
            if (q != k) {
                for (int r = tid; r < S; r += nt) {
                    double t;
                    t = Ar[r * S + k]; Ar[r * S + k] = Ar[r * S + q]; Ar[r * S + q] = t;
                    t = Ai[r * S + k]; Ai[r * S + k] = Ai[r * S + q]; Ai[r * S + q] = t;
                }
                __syncthreads();
            }
        }
        // left factor and diagonal right-hand side (replica_free_greens.h:477-496)
        for (int r = tid; r < S; r += nt) {
            const double x = a.beta * a.evals[r];
            const double sx = upper ? x : -x;
            const double c = fmax(0.0, sx);
            double le, re;
            if (upper) {
                le = (1.0 - tf) * x - (1.0 - theta) * c;
                re = tpf * x - theta * c;
            } else {
                le = tf * sx - (1.0 - theta) * c;
                re = (1.0 - tpf) * sx - theta * c;
            }
            colr[r] = upper ? exp(le) : -exp(le);
            coli[r] = exp(re);
        }
        __syncthreads();
        // T = V K with K[p][q] = (M^{-1}[p][q] e[q]) left[p], to scratch
        for (int e = tid; e < S2; e += nt) {
            const int i = e / S;
            const int q = e - i * S;
            const double eq = coli[q];
            double tr = 0.0, ti = 0.0;
            for (int p = 0; p < S; ++p) {
                const double v = a.V[i * S + p];
                const double kr = (Ar[p * S + q] * eq) * colr[p];
                const double ki = (Ai[p * S + q] * eq) * colr[p];
                tr = __builtin_fma(v, kr, tr);
                ti = __builtin_fma(v, ki, ti);
            }
            tr_g[e] = tr;
            ti_g[e] = ti;
        }
        __syncthreads();
        for (int e = tid; e < S2; e += nt) {
            Ar[e] = tr_g[e];
            Ai[e] = ti_g[e];
        }
        __syncthreads();
        // G = T V^T, to scratch
        for (int e = tid; e < S2; e += nt) {
            const int i = e / S;
            const int jx = e - i * S;
            double gr = 0.0, gi = 0.0;
            for (int q = 0; q < S; ++q) {
                const double v = a.V[jx * S + q];
                gr = __builtin_fma(Ar[i * S + q], v, gr);
                gi = __builtin_fma(Ai[i * S + q], v, gi);
            }
            tr_g[e] = gr;
            ti_g[e] = gi;
        }
        __syncthreads();
        const int64_t stride = corner_mode ? 1 : static_cast<int64_t>(a.nu) * a.nv;
        double* out = corner_mode ? a.corner : a.dst + node;
        if (a.ordinary) {
            for (int e = tid; e < S2; e += nt) out[static_cast<int64_t>(e) * stride] = tr_g[e];
        } else {
            const int SP = 4 * S2;
            for (int f = tid; f < SP; f += nt) {
                int g = f;
                const int y2 = g % a.N2;
                g /= a.N2;
                const int x2 = g % a.N1;
                g /= a.N1;
                const int y1 = g % a.N2;
                g /= a.N2;
                const int x1 = g % a.N1;
                g /= a.N1;
                const int rb = g & 1;
                const int ra = g >> 1;
                const int s1 = y1 * a.N1 + x1;
                const int s2 = y2 * a.N1 + x2;
                const double re = tr_g[s1 * S + s2];
                const double im = ti_g[s1 * S + s2];
                out[static_cast<int64_t>(f) * stride] = (ra == rb) ? re : (ra == 0 ? -im : im);
            }
        }
        // the scratch slot and the LDS are reused by the next node
        __syncthreads();
    }
}

size_t big_lds_bytes(int S) {
    return sizeof(double) * (2 * static_cast<size_t>(S) * S + 2 * S + kProdBigThreads / 64) +
           sizeof(int) * (kProdBigThreads / 64 + S + 1);
}

// In-place separable cosine transform of one sitepair's nu x nv sample grid.
__global__ __launch_bounds__(256) void k_dct(double* __restrict__ tab, int nu, int nv, const double* __restrict__ cos_u,
                                             const double* __restrict__ cos_v, double* __restrict__ gtmp) {
    extern __shared__ double tmp_lds[];
    const int64_t sp = blockIdx.x;
    double* blk = tab + sp * static_cast<int64_t>(nu) * nv;
    double* tmp = gtmp ? gtmp + sp * static_cast<int64_t>(nu) * nv : tmp_lds;
    const int tid = threadIdx.x;
    for (int e = tid; e < nu * nv; e += blockDim.x) {
        const int i = e / nv;
        const int b = e - i * nv;
        const double pv = (b == 0 ? 1.0 : 2.0) / static_cast<double>(nv);
        double t = 0.0;
        for (int j = 0; j < nv; ++j) t += blk[i * nv + j] * cos_v[b * nv + j];
        tmp[e] = pv * t;
    }
    __syncthreads();
    for (int e = tid; e < nu * nv; e += blockDim.x) {
        const int aa = e / nv;
        const int b = e - aa * nv;
        const double pu = (aa == 0 ? 1.0 : 2.0) / static_cast<double>(nu);
        double t = 0.0;
        for (int i = 0; i < nu; ++i) t += tmp[i * nv + b] * cos_u[aa * nu + i];
        blk[e] = pu * t;
    }
}

// The separable cosine transform as two FP64 matrix-core GEMMs (Nc a
// multiple of 16).  For every sitepair block X (nu x nv):
//   pass 1:  Y^T = (X Wv^T)^T          Wv[b][j] = (b ? 2 : 1)/nv cos(b theta_j)
//   pass 2:  C   = (Y^T Wu^T)^T = Wu X Wv^T
// Each pass is a row GEMM  out(r, n) = sum_k X[r][k] W[n][k]  over all rows r
// of the blocks, stored block-transposed: row r = blk * s + i goes to
// s * blk * N + n * blk + i.  A workgroup owns 64 rows; wave w computes the
// 16 x N tile out^T[n][r] = W X^T with W as the MFMA A operand and the rows as
// B, so that each accumulator register holds 16 consecutive rows of one n and
// the transposed store is contiguous.  K is staged through LDS 16 at a time.
typedef double pd4 __attribute__((ext_vector_type(4)));
constexpr int kDctRows = 64;
constexpr int kDctK = 16;

template <int NT>
__global__ __launch_bounds__(256) void k_dct_pass(const double* __restrict__ X, int64_t R, int K,
                                                  const double* __restrict__ W, int blk, double* __restrict__ out) {
    constexpr int N = NT * 16;
    __shared__ double xs[kDctRows][kDctK + 1];
    __shared__ double ws[N][kDctK + 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int g = lane >> 4;
    const int j = lane & 15;
    for (int64_t r0 = static_cast<int64_t>(blockIdx.x) * kDctRows; r0 < R;
         r0 += static_cast<int64_t>(gridDim.x) * kDctRows) {
        pd4 acc[NT];
        for (int t = 0; t < NT; ++t) acc[t] = pd4{0.0, 0.0, 0.0, 0.0};
        for (int kc = 0; kc < K; kc += kDctK) {
            for (int e = tid; e < kDctRows * kDctK; e += 256) {
                const int row = e / kDctK;
                const int kk = e - row * kDctK;
                const int64_t r = r0 + row;
                xs[row][kk] = r < R ? X[r * K + kc + kk] : 0.0;
            }
            for (int e = tid; e < N * kDctK; e += 256) {
                const int n = e / kDctK;
                const int kk = e - n * kDctK;
                ws[n][kk] = W[static_cast<int64_t>(n) * K + kc + kk];
            }
            __syncthreads();
#pragma unroll
            for (int ks = 0; ks < kDctK; ks += 4) {
                const double b = xs[wave * 16 + j][ks + g];
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const double a = ws[t * 16 + j][ks + g];
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
                }
            }
            __syncthreads();
        }
        const int64_t r = r0 + wave * 16 + j;
        if (r < R) {
            const int64_t blkno = r / blk;
            const int64_t base = blkno * blk * N + (r - blkno * blk);
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int q = 0; q < 4; ++q) out[base + static_cast<int64_t>(t * 16 + 4 * q + g) * blk] = acc[t][q];
        }
    }
}

template <int NT>
void launch_dct_pass_nt(const double* X, int64_t R, int K, const double* W, int blk, double* out, int grid,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_dct_pass<NT>, dim3(grid), dim3(256), 0, s, X, R, K, W, blk, out);
}

int launch_dct_pass(int N, const double* X, int64_t R, int K, const double* W, int blk, double* out, int ncu,
                    hipStream_t s) {
    const int64_t panels = (R + kDctRows - 1) / kDctRows;
    const int grid = static_cast<int>(std::min<int64_t>(panels, static_cast<int64_t>(ncu) * 8));
    switch (N / 16) {
#define CDET_DCT_NT(n) \
    case n: launch_dct_pass_nt<n>(X, R, K, W, blk, out, grid, s); break;
        CDET_DCT_NT(1) CDET_DCT_NT(2) CDET_DCT_NT(3) CDET_DCT_NT(4) CDET_DCT_NT(5) CDET_DCT_NT(6)
        CDET_DCT_NT(7) CDET_DCT_NT(8) CDET_DCT_NT(9) CDET_DCT_NT(10) CDET_DCT_NT(11) CDET_DCT_NT(12)
        CDET_DCT_NT(13) CDET_DCT_NT(14) CDET_DCT_NT(15) CDET_DCT_NT(16)
#undef CDET_DCT_NT
        default: return fail(CDET_ERR_CAPACITY, "cosine transform: Chebyshev count outside 16..256");
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// Cyclic Jacobi eigendecomposition of a real symmetric matrix (row-major),
// eigenvalues ascending, eigenvectors in columns of V.
void host_jacobi(std::vector<double> A, int n, std::vector<double>& ev, std::vector<double>& V) {
    V.assign(static_cast<size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) V[static_cast<size_t>(i) * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const double x = A[static_cast<size_t>(i) * n + j];
                tot += x * x;
                if (i != j) off += x * x;
            }
        if (off <= 1e-32 * std::max(tot, 1e-300)) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[static_cast<size_t>(p) * n + q];
                if (apq == 0.0) continue;
                const double app = A[static_cast<size_t>(p) * n + p];
                const double aqq = A[static_cast<size_t>(q) * n + q];
                const double th = (aqq - app) / (2.0 * apq);
                const double t = (th >= 0.0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0);
                const double s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[static_cast<size_t>(k) * n + p];
                    const double akq = A[static_cast<size_t>(k) * n + q];
                    A[static_cast<size_t>(k) * n + p] = c * akp - s * akq;
                    A[static_cast<size_t>(k) * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[static_cast<size_t>(p) * n + k];
                    const double aqk = A[static_cast<size_t>(q) * n + k];
                    A[static_cast<size_t>(p) * n + k] = c * apk - s * aqk;
                    A[static_cast<size_t>(q) * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[static_cast<size_t>(k) * n + p];
                    const double vkq = V[static_cast<size_t>(k) * n + q];
                    V[static_cast<size_t>(k) * n + p] = c * vkp - s * vkq;
                    V[static_cast<size_t>(k) * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> ord(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) ord[static_cast<size_t>(i)] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int l, int r) {
        return A[static_cast<size_t>(l) * n + l] < A[static_cast<size_t>(r) * n + r];
    });
    ev.assign(static_cast<size_t>(n), 0.0);
    std::vector<double> W(static_cast<size_t>(n) * n);
    for (int c = 0; c < n; ++c) {
        const int o = ord[static_cast<size_t>(c)];
        ev[static_cast<size_t>(c)] = A[static_cast<size_t>(o) * n + o];
        for (int r = 0; r < n; ++r) W[static_cast<size_t>(r) * n + c] = V[static_cast<size_t>(r) * n + o];
    }
    V.swap(W);
}

bool host_site_in_mask(int x, int y, int N1, int N2, int kind, int param) {
    switch (kind) {
        case 0: return false;
        case 1: return ((x + y) % 2) == 0;
        case 2: return y == 0;
        case 3: return x == 0;
        case 4: {
            auto in_iv = [&](int c, int extent) {
                if (param <= 0) return false;
                if (2 * param >= extent) return true;
                return c < param || c >= extent - param;
            };
            return in_iv(x, N1) && in_iv(y, N2);
        }
        case 5: return x < param;
        default: return false;
    }
}

// Device inputs of one table kind's producer: eigen-decomposition, sector
// mixing matrix, Chebyshev nodes and cosine tables.  `masked` applies the
// subsystem mask (twisted kinds); ordinary kinds use the empty mask.
struct Producer {
    int S = 0;
    double *d_ev = nullptr, *d_V = nullptr, *d_mre = nullptr, *d_mim = nullptr;
    double *d_nu = nullptr, *d_nv = nullptr, *d_cu = nullptr, *d_cv = nullptr;
    int32_t* d_err = nullptr;
    size_t lds = 0;
    bool big = false;          // k_node_green_big (S > 64)
    int grid = 0;              // its persistent grid
    double* d_tscr = nullptr;  // its scratch: grid slots of 2 S^2 doubles
    Producer() = default;
    Producer(const Producer&) = delete;
    Producer& operator=(const Producer&) = delete;
    ~Producer() {
        void* b[] = {d_ev, d_V, d_mre, d_mim, d_nu, d_nv, d_cu, d_cv, d_err, d_tscr};
        for (void* x : b)
            if (x) (void)hipFree(x);
    }
};

// Chebyshev nodes cos((2i+1)pi/2n) and cosine tables cos(a theta_i), [a][i]
// (build_chebyshev_node_tables, replica_chebyshev_tables.h:432-471), on the
// device.  Outputs are set before any allocation can fail, so a caller that
// frees non-null pointers cleans up after a partial failure.
int upload_node_tables(const cdet_params& p, double** d_nu, double** d_nv, double** d_cu, double** d_cv) {
    std::vector<double> nu_nodes(static_cast<size_t>(p.cheb_nu)), nv_nodes(static_cast<size_t>(p.cheb_nv));
    std::vector<double> cu(static_cast<size_t>(p.cheb_nu) * p.cheb_nu), cv(static_cast<size_t>(p.cheb_nv) * p.cheb_nv);
    const double pi = M_PI;
    for (int i = 0; i < p.cheb_nu; ++i) {
        const double th = (static_cast<double>(2 * i + 1) * pi) / static_cast<double>(2 * p.cheb_nu);
        nu_nodes[static_cast<size_t>(i)] = std::cos(th);
        for (int aa = 0; aa < p.cheb_nu; ++aa) cu[static_cast<size_t>(aa) * p.cheb_nu + i] = std::cos(static_cast<double>(aa) * th);
    }
    for (int j = 0; j < p.cheb_nv; ++j) {
        const double th = (static_cast<double>(2 * j + 1) * pi) / static_cast<double>(2 * p.cheb_nv);
        nv_nodes[static_cast<size_t>(j)] = std::cos(th);
        for (int b = 0; b < p.cheb_nv; ++b) cv[static_cast<size_t>(b) * p.cheb_nv + j] = std::cos(static_cast<double>(b) * th);
    }
    HIP_TRY(hipMalloc(d_nu, sizeof(double) * nu_nodes.size()));
    HIP_TRY(hipMalloc(d_nv, sizeof(double) * nv_nodes.size()));
    HIP_TRY(hipMalloc(d_cu, sizeof(double) * cu.size()));
    HIP_TRY(hipMalloc(d_cv, sizeof(double) * cv.size()));
    HIP_TRY(hipMemcpy(*d_nu, nu_nodes.data(), sizeof(double) * nu_nodes.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(*d_nv, nv_nodes.data(), sizeof(double) * nv_nodes.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(*d_cu, cu.data(), sizeof(double) * cu.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(*d_cv, cv.data(), sizeof(double) * cv.size(), hipMemcpyHostToDevice));
    return 0;
}

int producer_setup(cdet_ctx* ctx, double mu_eff, bool masked, const cdet_lattice* lat, Producer& pr) {
    const cdet_params& p = ctx->p;
    const int N1 = p.N1, N2 = p.N2, S = N1 * N2;
    if (S > kProdBigMaxSites)
        return fail(CDET_ERR_CAPACITY, "cdet_tables_build: lattices above 100 sites are not supported by this build");
    pr.S = S;
    // Square-lattice Hamiltonian (replica_free_greens.h:133-169).
    std::vector<double> H(static_cast<size_t>(S) * S, 0.0);
    for (int y = 0; y < N2; ++y)
        for (int x = 0; x < N1; ++x) {
            const int i = y * N1 + x;
            H[static_cast<size_t>(i) * S + i] = -mu_eff;
            int nb[2][2], nn = 0;
            if (x + 1 < N1) { nb[nn][0] = x + 1; nb[nn][1] = y; ++nn; }
            else if (lat->pbc) { nb[nn][0] = 0; nb[nn][1] = y; ++nn; }
            if (y + 1 < N2) { nb[nn][0] = x; nb[nn][1] = y + 1; ++nn; }
            else if (lat->pbc) { nb[nn][0] = x; nb[nn][1] = 0; ++nn; }
            for (int k = 0; k < nn; ++k) {
                const int j = nb[k][1] * N1 + nb[k][0];
                H[static_cast<size_t>(i) * S + j] -= 1.0;
                H[static_cast<size_t>(j) * S + i] -= 1.0;
            }
        }
    std::vector<double> ev, V;
    host_jacobi(H, S, ev, V);
    // S_0 = V^T Omega_0^{-1} V (Omega_0^{-1} = -i on the subsystem, 1 elsewhere).
    std::vector<double> mre(static_cast<size_t>(S) * S, 0.0), mim(static_cast<size_t>(S) * S, 0.0);
    for (int site = 0; site < S; ++site) {
        const bool in = masked && host_site_in_mask(site % N1, site / N1, N1, N2, lat->mask_kind, lat->mask_param);
        for (int a = 0; a < S; ++a)
            for (int b = 0; b < S; ++b) {
                const double w = V[static_cast<size_t>(site) * S + a] * V[static_cast<size_t>(site) * S + b];
                if (in)
                    mim[static_cast<size_t>(a) * S + b] -= w;
                else
                    mre[static_cast<size_t>(a) * S + b] += w;
            }
    }
    int rc = upload_node_tables(p, &pr.d_nu, &pr.d_nv, &pr.d_cu, &pr.d_cv);
    if (rc) return rc;
    HIP_TRY(hipMalloc(&pr.d_ev, sizeof(double) * S));
    HIP_TRY(hipMalloc(&pr.d_V, sizeof(double) * S * S));
    HIP_TRY(hipMalloc(&pr.d_mre, sizeof(double) * S * S));
    HIP_TRY(hipMalloc(&pr.d_mim, sizeof(double) * S * S));
    HIP_TRY(hipMalloc(&pr.d_err, sizeof(int32_t)));
    HIP_TRY(hipMemcpy(pr.d_ev, ev.data(), sizeof(double) * S, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(pr.d_V, V.data(), sizeof(double) * S * S, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(pr.d_mre, mre.data(), sizeof(double) * S * S, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(pr.d_mim, mim.data(), sizeof(double) * S * S, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(pr.d_err, 0, sizeof(int32_t)));
    if (S <= kProdMaxSites) {
        pr.lds = sizeof(double) * (4 * static_cast<size_t>(S) * S + 3 * S);
        if (pr.lds > 64 * 1024)
            HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_node_green),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(pr.lds)));
    } else {
        pr.big = true;
        pr.lds = big_lds_bytes(S);
        if (pr.lds > static_cast<size_t>(160 * 1024))
            return fail(CDET_ERR_CAPACITY, "cdet_tables_build: producer LDS exceeds 160 KiB");
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_node_green_big),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(pr.lds)));
        pr.grid = std::max(1, ctx->ncu);
        HIP_TRY(hipMalloc(&pr.d_tscr, sizeof(double) * 2 * static_cast<size_t>(S) * S * pr.grid));
    }
    return 0;
}

// Samples of one branch at [sitepair][node] in dst (nodes [node_begin,
// node_end)) and, when corner is not null, the corner values (tau = tau' = 0).
// replica00: the resident ordinary layout (replica (0,0), site-major).
int producer_samples(cdet_ctx* ctx, const Producer& pr, int upper, bool replica00, double* dst, double* corner,
                     int node_begin, int node_end, int node_chunk, hipStream_t s) {
    const cdet_params& p = ctx->p;
    ProdArgs a{};
    a.S = pr.S;
    a.N1 = p.N1;
    a.N2 = p.N2;
    a.nu = p.cheb_nu;
    a.nv = p.cheb_nv;
    a.ordinary = replica00 ? 1 : 0;
    a.upper = upper;
    a.beta = p.beta;
    a.evals = pr.d_ev;
    a.V = pr.d_V;
    a.mix_re = pr.d_mre;
    a.mix_im = pr.d_mim;
    a.node_u = pr.d_nu;
    a.node_v = pr.d_nv;
    a.dst = dst;
    a.corner = corner;
    a.err = pr.d_err;
    const int chunk = node_chunk > 0 ? node_chunk : node_end - node_begin;
    if (pr.big) {
        if (corner)
            hipLaunchKernelGGL(k_node_green_big, dim3(1), dim3(kProdBigThreads), pr.lds, s, a, 0, 1, 1, pr.d_tscr);
        for (int b0 = node_begin; b0 < node_end; b0 += chunk) {
            const int cnt = std::min(chunk, node_end - b0);
            hipLaunchKernelGGL(k_node_green_big, dim3(std::min(cnt, pr.grid)), dim3(kProdBigThreads), pr.lds, s, a,
                               b0, b0 + cnt, 0, pr.d_tscr);
        }
    } else {
        if (corner) hipLaunchKernelGGL(k_node_green, dim3(1), dim3(256), pr.lds, s, a, 0, 1);
        for (int b0 = node_begin; b0 < node_end; b0 += chunk) {
            const int cnt = std::min(chunk, node_end - b0);
            hipLaunchKernelGGL(k_node_green, dim3(cnt), dim3(256), pr.lds, s, a, b0, 0);
        }
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// In-place cosine transform of `active` sitepair sample grids at dst.  With
// nu and nv multiples of 16 (up to 256) it runs as the two matrix-core passes
// above, a bounded group of sitepairs at a time through a temporary of the
// same layout; otherwise the per-sitepair k_dct.  Returns with the stream
// drained.
int table_dct(cdet_ctx* ctx, const double* d_cu, const double* d_cv, double* dst, int64_t active, hipStream_t s) {
    const cdet_params& p = ctx->p;
    const int nu = p.cheb_nu, nv = p.cheb_nv;
    const size_t blk = static_cast<size_t>(nu) * nv;
    if (nu % 16 == 0 && nv % 16 == 0 && nu <= 256 && nv <= 256) {
        // scaled cosine tables W[a][i] = (a ? 2 : 1)/n cos(a theta_i)
        std::vector<double> wu(static_cast<size_t>(nu) * nu), wv(static_cast<size_t>(nv) * nv);
        const double pi = M_PI;
        for (int a = 0; a < nu; ++a)
            for (int i = 0; i < nu; ++i) {
                const double th = (static_cast<double>(2 * i + 1) * pi) / static_cast<double>(2 * nu);
                wu[static_cast<size_t>(a) * nu + i] =
                    (a == 0 ? 1.0 : 2.0) / static_cast<double>(nu) * std::cos(static_cast<double>(a) * th);
            }
        for (int b = 0; b < nv; ++b)
            for (int j = 0; j < nv; ++j) {
                const double th = (static_cast<double>(2 * j + 1) * pi) / static_cast<double>(2 * nv);
                wv[static_cast<size_t>(b) * nv + j] =
                    (b == 0 ? 1.0 : 2.0) / static_cast<double>(nv) * std::cos(static_cast<double>(b) * th);
            }
        const int64_t group = std::max<int64_t>(1, std::min<int64_t>(active, (int64_t{512} << 20) / (blk * 8)));
        double *d_w = nullptr, *tmp = nullptr;
        int rc = 0;
        if (hipMalloc(&d_w, sizeof(double) * (wu.size() + wv.size())) != hipSuccess ||
            hipMalloc(&tmp, sizeof(double) * blk * static_cast<size_t>(group)) != hipSuccess) {
            rc = fail(CDET_ERR_DEVICE, "cosine transform: device allocation failed");
        } else {
            double* d_wu = d_w;
            double* d_wv = d_w + wu.size();
            if (hipMemcpyAsync(d_wu, wu.data(), sizeof(double) * wu.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
                hipMemcpyAsync(d_wv, wv.data(), sizeof(double) * wv.size(), hipMemcpyHostToDevice, s) != hipSuccess)
                rc = fail(CDET_ERR_DEVICE, "cosine transform: table upload failed");
            for (int64_t sp0 = 0; sp0 < active && !rc; sp0 += group) {
                const int64_t cnt = std::min(group, active - sp0);
                double* x = dst + static_cast<size_t>(sp0) * blk;
                rc = launch_dct_pass(nv, x, cnt * nu, nv, d_wv, nu, tmp, ctx->ncu, s);
                if (!rc) rc = launch_dct_pass(nu, tmp, cnt * nv, nu, d_wu, nv, x, ctx->ncu, s);
            }
        }
        const hipError_t se = hipStreamSynchronize(s);
        if (d_w) (void)hipFree(d_w);
        if (tmp) (void)hipFree(tmp);
        if (rc) return rc;
        if (se != hipSuccess) return fail(CDET_ERR_DEVICE, std::string("cosine transform: ") + hipGetErrorString(se));
        return 0;
    }
    size_t dct_lds = blk * sizeof(double);
    double* gtmp = nullptr;
    if (dct_lds > kMaxLds) {
        HIP_TRY(hipMalloc(&gtmp, blk * sizeof(double) * static_cast<size_t>(active)));
        dct_lds = 0;
    } else if (dct_lds > 64 * 1024) {
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_dct), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(dct_lds)));
    }
    hipLaunchKernelGGL(k_dct, dim3(static_cast<unsigned>(active)), dim3(256), dct_lds, s, dst, nu, nv, d_cu, d_cv,
                       gtmp);
    const hipError_t le = hipGetLastError();
    const hipError_t se = hipStreamSynchronize(s);
    if (gtmp) (void)hipFree(gtmp);
    if (le != hipSuccess) return fail(CDET_ERR_DEVICE, std::string("k_dct: ") + hipGetErrorString(le));
    if (se != hipSuccess) return fail(CDET_ERR_DEVICE, std::string("k_dct: ") + hipGetErrorString(se));
    return 0;
}

int producer_status(const Producer& pr, hipStream_t s) {
    int32_t err = 0;
    HIP_TRY(hipMemcpyAsync(&err, pr.d_err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (err) return fail(CDET_ERR_SOLVE, "Complex matrix solve is singular within the requested precision");
    return 0;
}

// One table kind into the context's device table.  Twisted: mask as given,
// all replica blocks.  Ordinary: empty mask, either the resident replica-(0,0)
// site-major layout or the chunked full layout.
int build_one_table(cdet_ctx* ctx, int kind, double mu_eff, int ordinary, const cdet_lattice* lat) {
    const cdet_params& p = ctx->p;
    Producer pr;
    int rc = producer_setup(ctx, mu_eff, !ordinary, lat, pr);
    if (rc) return rc;
    rc = alloc_table(ctx, kind);
    if (rc) return rc;
    DevTable& t = ctx->tab[kind];
    invalidate_split(t);
    size_t ncoef, ncorner;
    table_sizes(ctx, ncoef, ncorner);
    hipStream_t s = ctx->stream;
    // cleared on the producer stream so that the clears are ordered before
    // the node kernels (the context stream does not synchronise with the
    // null stream)
    HIP_TRY(hipMemsetAsync(t.upper, 0, ncoef * sizeof(double), s));
    HIP_TRY(hipMemsetAsync(t.lower, 0, ncoef * sizeof(double), s));
    HIP_TRY(hipMemsetAsync(t.ucorner, 0, ncorner * sizeof(double), s));
    HIP_TRY(hipMemsetAsync(t.lcorner, 0, ncorner * sizeof(double), s));
    const bool replica00 = ordinary && lat->ordinary_layout == CDET_ORDINARY_RESIDENT;
    const int S = pr.S;
    const int64_t active = replica00 ? static_cast<int64_t>(S) * S : static_cast<int64_t>(4) * S * S;
    const int nodes = p.cheb_nu * p.cheb_nv;
    for (int br = 0; br < 2 && !rc; ++br) {
        double* dst = br == 0 ? t.upper : t.lower;
        rc = producer_samples(ctx, pr, br == 0 ? 1 : 0, replica00, dst, br == 0 ? t.ucorner : t.lcorner, 0, nodes,
                              lat->node_chunk, s);
        if (!rc) rc = table_dct(ctx, pr.d_cu, pr.d_cv, dst, active, s);
    }
    const int st = producer_status(pr, s);
    if (rc) return rc;
    if (st) return st;
    if (ordinary) ctx->ordinary_layout = lat->ordinary_layout;
    return 0;
}

}  // namespace

static int build_tables_gpu(cdet_ctx* ctx, const cdet_lattice* lattice) {
    if (!ctx || !lattice) return fail(CDET_ERR_ARGUMENT, "cdet_tables_build: bad argument");
    if (lattice->ordinary_layout != CDET_ORDINARY_RESIDENT && lattice->ordinary_layout != CDET_ORDINARY_CHUNKED)
        return fail(CDET_ERR_ARGUMENT, "cdet_tables_build: unknown ordinary_layout");
    const cdet_params& p = ctx->p;
    // Twisted tables use mu - alpha_sigma with the subsystem mask, ordinary
    // tables the same mu with the empty mask
    // (include/cdet/chunked_load_data.h:582-604).
    const bool equal = p.alpha_up == p.alpha_down;
    int rc = build_one_table(ctx, CDET_TABLE_TWISTED_UP, p.mu - p.alpha_up, 0, lattice);
    if (rc) return rc;
    if (equal) {
        rc = cdet_tables_alias(ctx, CDET_TABLE_TWISTED_DOWN, CDET_TABLE_TWISTED_UP);
    } else {
        rc = build_one_table(ctx, CDET_TABLE_TWISTED_DOWN, p.mu - p.alpha_down, 0, lattice);
    }
    if (rc) return rc;
    rc = build_one_table(ctx, CDET_TABLE_ORDINARY_UP, p.mu - p.alpha_up, 1, lattice);
    if (rc) return rc;
    if (equal) {
        rc = cdet_tables_alias(ctx, CDET_TABLE_ORDINARY_DOWN, CDET_TABLE_ORDINARY_UP);
    } else {
        rc = build_one_table(ctx, CDET_TABLE_ORDINARY_DOWN, p.mu - p.alpha_down, 1, lattice);
    }
    if (rc) return rc;
    refresh_scalars(ctx);
    return 0;
}

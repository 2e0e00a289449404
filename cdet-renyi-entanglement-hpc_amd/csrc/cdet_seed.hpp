// cdet_seed.hpp -- conditional generation probabilities of the histogram seed
// on the GPU (included by cdet_kernels.hip).
//
// Reference: compute_conditional_generation_probabilities
// (include/cdet/mcmcmc_sampler.h:252-330) with the factorized histogram
// density (include/cdet/mcmcmc_seed.h:130-150).  For an envelope of E
// vertices, with q(c|p) the seed density:
//   D_0[c]  = sum_p q(c|p)                           (p ascending, q(c|c) = 0)
//   D_G[c]  = D_{G minus lowest bit}[c] - q(c | lowest bit of G), values in
//             (-1e-12, 0) set to 0 after every subtraction
//   P[G]    = sum_{c in G, ascending} P[G \ c] D_G[c] / (|G| (E - |G|))
// with P[0] = 1 and P[full] = 0.  The host keeps D_G for every mask; here
// each thread forms D_G[c] for the c in its mask directly, by applying the
// bits of G from the highest to the lowest, which is the same sequence of
// subtractions and clamps.  One workgroup per envelope walks the popcount
// layers with P in LDS, so every value is the host's to the last bit.

namespace {

constexpr int kSeedThreads = 512;

struct SeedTables {
    const double* tau;    // bins x bins
    const double* space;  // S x S
    const double* rep;    // 2 x 2
    int bins, n1, n2, sites;
};

__device__ inline int seed_bin(double value, int extent) {
    const int i = static_cast<int>(floor(value * static_cast<double>(extent)));
    return i < 0 ? 0 : (i >= extent ? extent - 1 : i);
}

__device__ inline double seed_density(const SeedTables& t, const cdet_vertex& c, const cdet_vertex& p) {
    const int ps = seed_bin(p.x, t.n1) * t.n2 + seed_bin(p.y, t.n2);
    const int cs = seed_bin(c.x, t.n1) * t.n2 + seed_bin(c.y, t.n2);
    const double dt = t.tau[seed_bin(p.tau, t.bins) * t.bins + seed_bin(c.tau, t.bins)] * static_cast<double>(t.bins);
    const double ds = t.space[ps * t.sites + cs] * static_cast<double>(t.sites);
    const double dr = t.rep[seed_bin(p.replica, 2) * 2 + seed_bin(c.replica, 2)] * 2.0;
    return dt * ds * dr;
}

__global__ __launch_bounds__(kSeedThreads) void k_seed_conditional(SeedTables t, const cdet_vertex* __restrict__ verts,
                                                                   const int64_t* __restrict__ vbase,
                                                                   const int32_t* __restrict__ orders,
                                                                   const int64_t* __restrict__ obase,
                                                                   const uint16_t* __restrict__ masks,
                                                                   const int* __restrict__ popoff,
                                                                   double* __restrict__ out, int32_t* __restrict__ err) {
    extern __shared__ double sp[];
    const int env = blockIdx.x;
    const int E = orders[env];
    const int full = (1 << E) - 1;
    double* P = sp;                      // 2^E
    double* dens = P + (full + 1);       // E x E, [child][parent]
    const cdet_vertex* v = verts + vbase[env];
    const int tid = threadIdx.x;
    for (int e = tid; e < E * E; e += blockDim.x) {
        const int c = e / E;
        const int p = e - c * E;
        double d = 0.0;
        if (c != p) {
            d = seed_density(t, v[c], v[p]);
            if (d < 0.0 || !isfinite(d)) atomicOr(err, 1);
        }
        dens[e] = d;
    }
    for (int g = tid; g <= full; g += blockDim.x) P[g] = g == 0 ? 1.0 : 0.0;
    __syncthreads();
    const uint16_t* ms = masks + static_cast<size_t>(E) * (1u << kMaxOrder);
    const int* off = popoff + E * (kMaxOrder + 2);
    for (int m = 1; m < E; ++m) {
        for (int w = off[m] + tid; w < off[m + 1]; w += blockDim.x) {
            const int g = ms[w];
            double s = 0.0;
            for (int c = 0; c < E; ++c) {
                if (!((g >> c) & 1)) continue;
                double d = 0.0;
                for (int p = 0; p < E; ++p) d += dens[c * E + p];
                for (int b = E - 1; b >= 0; --b) {
                    if (!((g >> b) & 1)) continue;
                    d = d - dens[c * E + b];
                    if (d < 0.0 && d > -1.0e-12) d = 0.0;
                }
                s += P[g ^ (1 << c)] * d;
            }
            P[g] = s / (static_cast<double>(m) * static_cast<double>(E - m));
        }
        __syncthreads();
    }
    double* o = out + obase[env];
    for (int g = tid; g <= full; g += blockDim.x) o[g] = g == full ? 0.0 : P[g];
}

uint64_t seed_hash(const cdet_histogram_seed& h) {
    uint64_t x = 1469598103934665603ULL;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) x = (x ^ b[i]) * 1099511628211ULL;
    };
    mix(&h.tau_bins, sizeof(int32_t) * 3);
    const size_t S = static_cast<size_t>(h.N1) * h.N2;
    mix(h.tau_prob, sizeof(double) * static_cast<size_t>(h.tau_bins) * h.tau_bins);
    mix(h.space_prob, sizeof(double) * S * S);
    mix(h.replica_prob, sizeof(double) * 4);
    return x;
}

}  // namespace

// GPU form of the per-step conditional generation probabilities for a batch
// of envelopes (host buffers in and out, as cdet_eval_batch_host).
extern "C" int cdet_seed_cond_batch_host(void* user, const cdet_histogram_seed* seed, const cdet_vertex* h_vertices,
                                         const int32_t* h_orders, int64_t count, double* h_out) {
    cdet_ctx* ctx = static_cast<cdet_ctx*>(user);
    if (!ctx || !seed || !seed->tau_prob || !seed->space_prob || !h_vertices || !h_orders || !h_out || count <= 0)
        return fail(CDET_ERR_ARGUMENT, "cdet_seed_cond_batch_host: bad argument");
    if (seed->tau_bins <= 0 || seed->N1 <= 0 || seed->N2 <= 0)
        return fail(CDET_ERR_ARGUMENT, "Histogram seed dimensions must be positive");
    std::vector<int64_t> vb(static_cast<size_t>(count)), ob(static_cast<size_t>(count) + 1, 0);
    int64_t nv = 0;
    int emax = 1;
    for (int64_t i = 0; i < count; ++i) {
        const int e = h_orders[i];
        if (e < 1 || e > kMaxOrder) return fail(CDET_ERR_CAPACITY, "envelope order above CDET_MAX_ORDER");
        vb[static_cast<size_t>(i)] = nv;
        nv += e;
        ob[static_cast<size_t>(i) + 1] = ob[static_cast<size_t>(i)] + (int64_t{1} << e);
        emax = std::max(emax, e);
    }
    const size_t S = static_cast<size_t>(seed->N1) * seed->N2;
    const size_t tab_doubles = static_cast<size_t>(seed->tau_bins) * seed->tau_bins + S * S + 4;
    const uint64_t h = seed_hash(*seed);
    int rc = 0;
    if (h != ctx->seed_hash || ctx->seed_tab_cap < tab_doubles * sizeof(double)) {
        rc = grow(&ctx->d_seed_tab, ctx->seed_tab_cap, tab_doubles * sizeof(double));
        if (rc) return rc;
        std::vector<double> packed(tab_doubles);
        const size_t nt = static_cast<size_t>(seed->tau_bins) * seed->tau_bins;
        std::copy(seed->tau_prob, seed->tau_prob + nt, packed.begin());
        std::copy(seed->space_prob, seed->space_prob + S * S, packed.begin() + static_cast<std::ptrdiff_t>(nt));
        std::copy(seed->replica_prob, seed->replica_prob + 4, packed.begin() + static_cast<std::ptrdiff_t>(nt + S * S));
        HIP_TRY(hipMemcpy(ctx->d_seed_tab, packed.data(), tab_doubles * sizeof(double), hipMemcpyHostToDevice));
        ctx->seed_hash = h;
    }
    const size_t vbytes = sizeof(cdet_vertex) * static_cast<size_t>(nv);
    const size_t meta = sizeof(int64_t) * (2 * static_cast<size_t>(count) + 1) + sizeof(int32_t) * (count + 1);
    const size_t obytes = sizeof(double) * static_cast<size_t>(ob.back());
    rc = grow(&ctx->d_seed_in, ctx->seed_in_cap, vbytes + meta + 64);
    if (!rc) rc = grow(&ctx->d_seed_out, ctx->seed_out_cap, obytes);
    if (rc) return rc;
    char* base = static_cast<char*>(ctx->d_seed_in);
    cdet_vertex* d_v = reinterpret_cast<cdet_vertex*>(base);
    int64_t* d_vb = reinterpret_cast<int64_t*>(base + ((vbytes + 15) & ~size_t{15}));
    int64_t* d_ob = d_vb + count;
    int32_t* d_ord = reinterpret_cast<int32_t*>(d_ob + count + 1);
    int32_t* d_err = d_ord + count;
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(d_v, h_vertices, vbytes, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_vb, vb.data(), sizeof(int64_t) * count, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_ob, ob.data(), sizeof(int64_t) * (count + 1), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_ord, h_orders, sizeof(int32_t) * count, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, d_err, static_cast<int64_t>(1), 0);
    const double* tab = static_cast<const double*>(ctx->d_seed_tab);
    const size_t ntau = static_cast<size_t>(seed->tau_bins) * seed->tau_bins;
    SeedTables t{tab, tab + ntau, tab + ntau + S * S, seed->tau_bins, seed->N1, seed->N2, static_cast<int>(S)};
    const size_t lds = sizeof(double) * ((size_t{1} << emax) + static_cast<size_t>(emax) * emax);
    if (lds > 64 * 1024)
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_seed_conditional),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    hipLaunchKernelGGL(k_seed_conditional, dim3(static_cast<unsigned>(count)), dim3(kSeedThreads), lds, s, t, d_v, d_vb,
                       d_ord, d_ob, ctx->d_masks, ctx->d_popoff, static_cast<double*>(ctx->d_seed_out), d_err);
    HIP_TRY(hipGetLastError());
    int32_t err = 0;
    HIP_TRY(hipMemcpyAsync(h_out, ctx->d_seed_out, obytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&err, d_err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (err) return fail(CDET_ERR_NONFINITE, "MCMCMC seed density is negative or non-finite");
    return 0;
}

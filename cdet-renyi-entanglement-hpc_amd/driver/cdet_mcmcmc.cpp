// cdet_mcmcmc.cpp -- lockstep set heat-bath MCMCMC driver (host).
//
// Reproduces the reference sampler chain by chain
// (include/cdet/mcmcmc_sampler.h:159-1145, include/cdet/mcmcmc_seed.h):
// every chain keeps its own std::mt19937_64 and consumes it in the
// reference's order (envelope draws, uniform_real_distribution(0, W) for the
// subset choice, the parent draw of the seed regeneration, then the seed
// kernel's draws per regenerated vertex: four uniforms for the uniform seed,
// three table draws and four uniforms for the histogram seed), so its
// trajectory and checkpoint bytes equal those of the reference's per-chain
// thread.  What changes is the schedule: all live
// chains advance one step together and their envelopes are evaluated by one
// batched integrand call.
//
// Built with g++/libstdc++ (the reference's standard library) so that the
// distributions and the mt19937_64 stream text are the reference's.

#include "../../include/cdet_mcmcmc.h"

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <limits>
#include <map>
#include <mutex>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_error;

struct Stats {  // RunningStatsLongDouble (mcmcmc_sampler.h:31-78)
    uint64_t count = 0;
    long double mean = 0.0L;
    long double m2 = 0.0L;
    void add(long double v) {
        ++count;
        const long double d = v - mean;
        mean += d / static_cast<long double>(count);
        const long double d2 = v - mean;
        m2 += d * d2;
    }
    long double variance() const {
        return count < 2 ? std::numeric_limits<long double>::quiet_NaN() : m2 / static_cast<long double>(count - 1);
    }
    long double stderr_() const {
        return count < 2 ? std::numeric_limits<long double>::quiet_NaN()
                         : std::sqrt(variance() / static_cast<long double>(count));
    }
};

uint64_t splitmix64(uint64_t x) {  // mcmcmc_sampler.h:159-164
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30U)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27U)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31U);
}

uint64_t derive_seed(uint64_t base, uint64_t tag, uint64_t index) {  // :166-172
    uint64_t mixed = splitmix64(base ^ tag);
    mixed ^= 0x9e3779b97f4a7c15ULL * (index + 1ULL);
    return splitmix64(mixed);
}

double binomial_as_double(int n, int k) {  // :174-185
    if (k < 0 || k > n) return 0.0;
    k = std::min(k, n - k);
    long double r = 1.0L;
    for (int i = 1; i <= k; ++i) {
        r *= static_cast<long double>(n - k + i);
        r /= static_cast<long double>(i);
    }
    return static_cast<double>(r);
}

cdet_vertex uniform_vertex(std::mt19937_64& rng) {  // :187-190 (braced order tau, x, y, replica)
    static thread_local std::uniform_real_distribution<double> u01(0.0, 1.0);
    const double t = u01(rng);
    const double x = u01(rng);
    const double y = u01(rng);
    const double r = u01(rng);
    return cdet_vertex{t, x, y, r};
}

struct MaskPlan {  // :403-431
    uint64_t subset_count = 0, full_mask = 0;
    std::vector<uint64_t> masks;
    std::vector<int> orders;
    std::vector<double> inv_binom;
};

constexpr int kMaxEnvelope = 48;  // highest envelope order (validate)

MaskPlan make_plan(const cdet_mcmcmc_config& c) {
    MaskPlan p;
    p.subset_count = uint64_t{1} << c.envelope_order;
    p.full_mask = p.subset_count - 1U;
    p.inv_binom.assign(static_cast<size_t>(c.envelope_order + 1), 0.0);
    for (int o = c.order_min; o <= c.order_max; ++o)
        p.inv_binom[static_cast<size_t>(o)] = 1.0 / binomial_as_double(c.envelope_order, o);
    for (uint64_t m = 1; m < p.subset_count; ++m) {
        const int o = __builtin_popcountll(m);
        if (o < c.order_min || o > c.order_max) continue;
        p.masks.push_back(m);
        p.orders.push_back(o);
    }
    return p;
}

void validate(const cdet_mcmcmc_config& c) {  // :465-521
    if (c.order_min < 1) throw std::runtime_error("MCMCMC_ORDER_MIN must be at least 1 for the replica integrand");
    if (c.order_max < c.order_min) throw std::runtime_error("MCMCMC_ORDER_MAX must be >= MCMCMC_ORDER_MIN");
    if (c.envelope_order <= c.order_max)
        throw std::runtime_error("MCMCMC_ENVELOPE_ORDER must be larger than MCMCMC_ORDER_MAX");
    if (c.envelope_order > kMaxEnvelope)
        throw std::runtime_error("MCMCMC_ENVELOPE_ORDER must be <= 48 for the current principal-minor backend");
    if (c.normalization_order < c.order_min || c.normalization_order > c.order_max)
        throw std::runtime_error("MCMCMC_NORMALIZATION_ORDER must lie inside the measured order range");
    if (c.chain_count <= 0) throw std::runtime_error("MCMCMC_CHAIN_COUNT must be positive");
    if (c.lambda_warmup_steps < 0 || c.production_steps <= 0)
        throw std::runtime_error("MCMCMC warmup must be nonnegative and production must be positive");
    if (c.lambda_adapt_interval <= 0) throw std::runtime_error("MCMCMC_LAMBDA_ADAPT_INTERVAL must be positive");
    if (c.seed_mode != CDET_SEED_UNIFORM && c.seed_mode != CDET_SEED_HISTOGRAM)
        throw std::runtime_error("MCMCMC_SEED_MODE must be uniform or histogram");
    if (c.resume_mode < CDET_RESUME_OFF || c.resume_mode > CDET_RESUME_REQUIRE)
        throw std::runtime_error("MCMCMC_RESUME must be off, auto, or require");
    if (c.checkpoint_enabled && c.checkpoint_interval_steps == 0)
        throw std::runtime_error("MCMCMC_CHECKPOINT_INTERVAL_STEPS must be positive when checkpointing is enabled");
    if (c.seed_mode == CDET_SEED_HISTOGRAM && c.histogram_learning_steps <= 0)
        throw std::runtime_error("MCMCMC_SEED_HISTOGRAM_LEARNING_STEPS must be positive for histogram seed");
    if (c.histogram_tau_bins <= 0) throw std::runtime_error("MCMCMC_SEED_HISTOGRAM_TAU_BINS must be positive");
    if (c.histogram_uniform_mix < 0.0 || c.histogram_uniform_mix > 1.0)
        throw std::runtime_error("MCMCMC_SEED_HISTOGRAM_UNIFORM_MIX must lie in [0, 1]");
    if (c.histogram_pseudo_count < 0.0 || !std::isfinite(c.histogram_pseudo_count))
        throw std::runtime_error("MCMCMC_SEED_HISTOGRAM_PSEUDO_COUNT must be nonnegative and finite");
    for (int o = c.order_min; o <= c.order_max; ++o) {
        const double l = c.lambdas[o];
        if (!(l > 0.0) || !std::isfinite(l)) throw std::runtime_error("MCMCMC_LAMBDAS entries must be positive finite numbers");
    }
}

const char* seed_mode_name(int m) { return m == CDET_SEED_HISTOGRAM ? "histogram" : "uniform"; }

// ---- seed kernels (include/cdet/mcmcmc_seed.h) -----------------------------

// Bin of a unit coordinate: floor(value * extent) clamped to [0, extent - 1]
// (unit_coordinate_to_index, mcmcmc_seed.h:30-39).
int unit_bin(double value, int extent) {
    const int i = static_cast<int>(std::floor(value * static_cast<double>(extent)));
    return i < 0 ? 0 : (i >= extent ? extent - 1 : i);
}

// The uniform seed (mcmcmc_seed.h:41-61) when `uniform`, else the factorized
// histogram seed (:71-229).  Tables are [parent][child], row-major; each CDF
// row is the running sum of its probabilities with the last entry forced to 1.
struct SeedKernel {
    bool uniform = true;
    int bins = 1, n1 = 1, n2 = 1, sites = 1;
    double mix = 0.0, pseudo = 0.0;
    std::vector<double> tau, space, rep;
    std::vector<double> tau_cdf, space_cdf, rep_cdf;

    static void make_cdf(const std::vector<double>& p, int cols, std::vector<double>& cdf) {
        cdf.assign(p.size(), 0.0);
        const size_t rows = p.size() / static_cast<size_t>(cols);
        for (size_t r = 0; r < rows; ++r) {
            double run = 0.0;
            const size_t o = r * static_cast<size_t>(cols);
            for (int k = 0; k < cols; ++k) {
                run += p[o + static_cast<size_t>(k)];
                cdf[o + static_cast<size_t>(k)] = run;
            }
            cdf[o + static_cast<size_t>(cols - 1)] = 1.0;
        }
    }

    static SeedKernel from(const cdet_histogram_seed* h) {
        SeedKernel s;
        if (!h) return s;
        if (h->tau_bins <= 0 || h->N1 <= 0 || h->N2 <= 0)
            throw std::runtime_error("Histogram seed dimensions must be positive");
        if (!h->tau_prob || !h->space_prob) throw std::runtime_error("histogram seed tables are null");
        s.uniform = false;
        s.bins = h->tau_bins;
        s.n1 = h->N1;
        s.n2 = h->N2;
        s.sites = h->N1 * h->N2;
        s.mix = h->uniform_mix;
        s.pseudo = h->pseudo_count;
        s.tau.assign(h->tau_prob, h->tau_prob + static_cast<size_t>(s.bins) * static_cast<size_t>(s.bins));
        s.space.assign(h->space_prob, h->space_prob + static_cast<size_t>(s.sites) * static_cast<size_t>(s.sites));
        s.rep.assign(h->replica_prob, h->replica_prob + 4);
        make_cdf(s.tau, s.bins, s.tau_cdf);
        make_cdf(s.space, s.sites, s.space_cdf);
        make_cdf(s.rep, 2, s.rep_cdf);
        return s;
    }

    int site_of(const cdet_vertex& v) const { return unit_bin(v.x, n1) * n2 + unit_bin(v.y, n2); }

    // One categorical draw from CDF row `row`: the first column whose CDF is
    // not below a uniform threshold (mcmcmc_seed.h:206-219).
    static int draw(const std::vector<double>& cdf, int cols, int row, std::mt19937_64& rng) {
        static thread_local std::uniform_real_distribution<double> u01(0.0, 1.0);
        const double threshold = u01(rng);
        const auto first = cdf.begin() + static_cast<std::ptrdiff_t>(row) * cols;
        const auto last = first + cols;
        const auto it = std::lower_bound(first, last, threshold);
        return it == last ? cols - 1 : static_cast<int>(it - first);
    }

    // sample_child_from_parent: time bin, site and replica draws, then one
    // uniform offset per coordinate in the order tau, x, y, replica.
    cdet_vertex child(const cdet_vertex& parent, std::mt19937_64& rng) const {
        if (uniform) return uniform_vertex(rng);
        static thread_local std::uniform_real_distribution<double> u01(0.0, 1.0);
        const int ct = draw(tau_cdf, bins, unit_bin(parent.tau, bins), rng);
        const int cs = draw(space_cdf, sites, site_of(parent), rng);
        const int cx = cs / n2;
        const int cy = cs - cx * n2;
        const int cr = draw(rep_cdf, 2, unit_bin(parent.replica, 2), rng);
        const double t = (static_cast<double>(ct) + u01(rng)) / static_cast<double>(bins);
        const double x = (static_cast<double>(cx) + u01(rng)) / static_cast<double>(n1);
        const double y = (static_cast<double>(cy) + u01(rng)) / static_cast<double>(n2);
        const double r = (static_cast<double>(cr) + u01(rng)) / 2.0;
        return cdet_vertex{t, x, y, r};
    }

    // density_child_given_parent: product of the three table entries, each
    // scaled by its number of columns (a density on the unit cube).
    double density(const cdet_vertex& c, const cdet_vertex& p) const {
        if (uniform) return 1.0;
        const double dt = tau[static_cast<size_t>(unit_bin(p.tau, bins)) * static_cast<size_t>(bins) +
                              static_cast<size_t>(unit_bin(c.tau, bins))] *
                          static_cast<double>(bins);
        const double ds = space[static_cast<size_t>(site_of(p)) * static_cast<size_t>(sites) +
                                static_cast<size_t>(site_of(c))] *
                          static_cast<double>(sites);
        const double dr = rep[static_cast<size_t>(unit_bin(p.replica, 2)) * 2 + static_cast<size_t>(unit_bin(c.replica, 2))] * 2.0;
        return dt * ds * dr;
    }
};

// Conditional generation probabilities (mcmcmc_sampler.h:252-330, double
// recursion).  out[G] is the density of regenerating the vertex subset G from
// the kept set (all vertices outside G): with q(c|p) the seed density,
//   D[G][c] = sum_{p not in G} q(c|p),
//   out[G]  = sum_{c in G} out[G \ c] * D[G][c] / (|G| * (E - |G|)),
// out[0] = 1, out[full] = 0.  D[G] is formed from D[G minus its lowest bit]
// by one subtraction per child, with results in (-1e-12, 0) set to zero.
// `dens` and `comp` are caller workspaces.
void conditional_probabilities(const SeedKernel& seed, const cdet_vertex* v, int E, std::vector<double>& out,
                               std::vector<double>& dens, std::vector<double>& comp) {
    const uint64_t count = uint64_t{1} << E;
    const size_t e = static_cast<size_t>(E);
    if (seed.uniform) {
        out.assign(count, 1.0);
        return;
    }
    dens.assign(e * e, 0.0);
    for (int c = 0; c < E; ++c)
        for (int p = 0; p < E; ++p) {
            if (c == p) continue;
            const double d = seed.density(v[c], v[p]);
            if (d < 0.0 || !std::isfinite(d)) throw std::runtime_error("MCMCMC seed density is negative or non-finite");
            dens[static_cast<size_t>(c) * e + static_cast<size_t>(p)] = d;
        }
    comp.assign(count * e, 0.0);
    for (int c = 0; c < E; ++c) {
        double s = 0.0;
        for (int p = 0; p < E; ++p) s += dens[static_cast<size_t>(c) * e + static_cast<size_t>(p)];
        comp[static_cast<size_t>(c)] = s;
    }
    out.assign(count, 0.0);
    out[0] = 1.0;
    const uint64_t full = count - 1U;
    for (uint64_t g = 1; g < full; ++g) {
        const uint64_t low = g & (~g + 1U);
        const size_t rp = static_cast<size_t>(__builtin_ctzll(low));
        const double* prev = comp.data() + static_cast<size_t>(g ^ low) * e;
        double* cur = comp.data() + static_cast<size_t>(g) * e;
        for (size_t c = 0; c < e; ++c) {
            double d = prev[c] - dens[c * e + rp];
            if (d < 0.0 && d > -1.0e-12) d = 0.0;
            cur[c] = d;
        }
        const int gen = __builtin_popcountll(g);
        double s = 0.0;
        for (size_t c = 0; c < e; ++c) {
            const uint64_t bit = uint64_t{1} << c;
            if (!(g & bit)) continue;
            s += out[static_cast<size_t>(g ^ bit)] * cur[c];
        }
        out[static_cast<size_t>(g)] = s / (static_cast<double>(gen) * static_cast<double>(E - gen));
    }
    out[static_cast<size_t>(full)] = 0.0;
}

// Learning accumulator (FactorizedHistogramSeedAccumulator,
// mcmcmc_seed.h:231-337): long-double directed-pair weights per table.
struct HistogramCounts {
    int bins, n1, n2, sites;
    double mix, pseudo;
    std::vector<long double> tau, space, rep;
    long double total = 0.0L;

    HistogramCounts(int b, int x, int y, double m, double ps) : bins(b), n1(x), n2(y), sites(x * y), mix(m), pseudo(ps) {
        if (bins <= 0 || n1 <= 0 || n2 <= 0) throw std::runtime_error("Histogram seed dimensions must be positive");
        if (mix < 0.0 || mix > 1.0) throw std::runtime_error("MCMCMC_SEED_HISTOGRAM_UNIFORM_MIX must lie in [0, 1]");
        if (pseudo < 0.0 || !std::isfinite(pseudo))
            throw std::runtime_error("MCMCMC_SEED_HISTOGRAM_PSEUDO_COUNT must be nonnegative and finite");
        tau.assign(static_cast<size_t>(bins) * static_cast<size_t>(bins), 0.0L);
        space.assign(static_cast<size_t>(sites) * static_cast<size_t>(sites), 0.0L);
        rep.assign(4, 0.0L);
    }
    int site_of(const cdet_vertex& v) const { return unit_bin(v.x, n1) * n2 + unit_bin(v.y, n2); }

    void add(const cdet_vertex& p, const cdet_vertex& c, long double w) {
        if (!(w > 0.0L) || !std::isfinite(static_cast<double>(w))) return;
        tau[static_cast<size_t>(unit_bin(p.tau, bins)) * static_cast<size_t>(bins) + static_cast<size_t>(unit_bin(c.tau, bins))] += w;
        space[static_cast<size_t>(site_of(p)) * static_cast<size_t>(sites) + static_cast<size_t>(site_of(c))] += w;
        rep[static_cast<size_t>(unit_bin(p.replica, 2)) * 2 + static_cast<size_t>(unit_bin(c.replica, 2))] += w;
        total += w;
    }

    // Row normalisation with pseudo-counts and uniform mixing in long double,
    // then a renormalisation of each row in double (normalize_rows, :292-329).
    void normalize(const std::vector<long double>& cnt, int cols, double* out) const {
        const size_t rows = cnt.size() / static_cast<size_t>(cols);
        const long double uni = 1.0L / static_cast<long double>(cols);
        const long double m = static_cast<long double>(mix);
        const long double ps = static_cast<long double>(pseudo);
        for (size_t r = 0; r < rows; ++r) {
            const long double* row = cnt.data() + r * static_cast<size_t>(cols);
            double* o = out + r * static_cast<size_t>(cols);
            long double sum = 0.0L;
            for (int k = 0; k < cols; ++k) sum += row[k];
            const long double den = sum + ps * static_cast<long double>(cols);
            for (int k = 0; k < cols; ++k) {
                const long double learned = den > 0.0L ? (row[k] + ps) / den : uni;
                o[k] = static_cast<double>((1.0L - m) * learned + m * uni);
            }
            double t = 0.0;
            for (int k = 0; k < cols; ++k) t += o[k];
            if (t > 0.0 && std::isfinite(t))
                for (int k = 0; k < cols; ++k) o[k] /= t;
        }
    }
};

// learn_factorized_histogram_seed (mcmcmc_sampler.h:607-671).  The envelopes
// come from one stream, in order, so they are drawn a chunk at a time and
// evaluated as one batch; the pair weights are then accumulated envelope by
// envelope, mask by mask, parent by parent, child by child, which is the
// reference's summation order for every long-double bin.
void learn_histogram(const cdet_mcmcmc_config& c, cdet_batch_eval_fn eval, void* user, cdet_histogram_seed& h,
                     double* total_out) {
    HistogramCounts acc(c.histogram_tau_bins, c.N1, c.N2, c.histogram_uniform_mix, c.histogram_pseudo_count);
    constexpr uint64_t kHistogramTag = 0x484953544F534545ULL;  // :618
    std::mt19937_64 rng(derive_seed(c.base_seed, kHistogramTag, 0));
    const int E = c.envelope_order;
    const size_t sc = size_t{1} << E;
    std::vector<double> pair_w(static_cast<size_t>(E + 1), 0.0);
    std::vector<long double> pair_wl(static_cast<size_t>(E + 1), 0.0L);
    // Envelopes per batch: bounded by 64 MiB of integrand values.
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(8192, (int64_t{64} << 20) / static_cast<int64_t>(sc * 8)));
    std::vector<cdet_vertex> hv;
    std::vector<int32_t> ho;
    std::vector<double> hout;
    for (int64_t done = 0; done < c.histogram_learning_steps;) {
        const int64_t n = std::min<int64_t>(chunk, c.histogram_learning_steps - done);
        hv.resize(static_cast<size_t>(n * E));
        for (auto& v : hv) v = uniform_vertex(rng);
        ho.assign(static_cast<size_t>(n), E);
        hout.assign(static_cast<size_t>(n) * sc, 0.0);
        const int rc = eval(user, hv.data(), ho.data(), n, hout.data());
        if (rc != 0) throw std::runtime_error("integrand evaluation failed (status " + std::to_string(rc) + ")");
        for (int64_t k = 0; k < n; ++k) {
            const cdet_vertex* v = hv.data() + k * E;
            const double* val = hout.data() + static_cast<size_t>(k) * sc;
            for (uint64_t mask = 1; mask < sc; ++mask) {
                const int order = __builtin_popcountll(mask);
                if (order < c.order_min || order > c.order_max || order < 2) continue;
                const double x = val[mask];
                if (!std::isfinite(x)) continue;
                const long double subset_w = static_cast<long double>(std::abs(x)) /
                                             static_cast<long double>(binomial_as_double(E, order));
                const long double w = subset_w / static_cast<long double>(order * (order - 1));
                if (!(w > 0.0L)) continue;
                for (int p = 0; p < E; ++p) {
                    if (!(mask & (uint64_t{1} << p))) continue;
                    for (int ch = 0; ch < E; ++ch) {
                        if (ch == p || !(mask & (uint64_t{1} << ch))) continue;
                        acc.add(v[p], v[ch], w);
                    }
                }
            }
        }
        done += n;
    }
    h.tau_bins = acc.bins;
    h.N1 = acc.n1;
    h.N2 = acc.n2;
    h.uniform_mix = acc.mix;
    h.pseudo_count = acc.pseudo;
    acc.normalize(acc.tau, acc.bins, h.tau_prob);
    acc.normalize(acc.space, acc.sites, h.space_prob);
    acc.normalize(acc.rep, 2, h.replica_prob);
    if (total_out) *total_out = static_cast<double>(acc.total);
}

void write_matrix_csv(const std::string& path, const double* v, int rows, int cols) {  // src/main.cpp:704-723
    std::ofstream out(path);
    if (!out.is_open()) throw std::runtime_error("Could not open histogram output file: " + path);
    out << std::setprecision(17);
    for (int r = 0; r < rows; ++r) {
        for (int k = 0; k < cols; ++k) {
            if (k > 0) out << ',';
            out << v[static_cast<size_t>(r) * static_cast<size_t>(cols) + static_cast<size_t>(k)];
        }
        out << '\n';
    }
}

void export_histogram(const cdet_histogram_seed& h, double beta, const std::string& prefix) {  // :725-767
    if (prefix.empty()) return;
    const std::filesystem::path parent = std::filesystem::path(prefix).parent_path();
    if (!parent.empty()) std::filesystem::create_directories(parent);
    const int sites = h.N1 * h.N2;
    write_matrix_csv(prefix + "_tau.csv", h.tau_prob, h.tau_bins, h.tau_bins);
    write_matrix_csv(prefix + "_space.csv", h.space_prob, sites, sites);
    write_matrix_csv(prefix + "_replica.csv", h.replica_prob, 2, 2);
    std::ofstream meta(prefix + "_metadata.txt");
    if (!meta.is_open()) throw std::runtime_error("Could not open histogram metadata file");
    meta << "N1=" << h.N1 << '\n'
         << "N2=" << h.N2 << '\n'
         << "BETA=" << beta << '\n'
         << "TAU_BINS=" << h.tau_bins << '\n'
         << "UNIFORM_MIX=" << h.uniform_mix << '\n'
         << "PSEUDO_COUNT=" << h.pseudo_count << '\n';
}

// Fixed pool of host threads for the per-chain step work.  run(n, f) calls
// f(0..n-1), each index once, on the caller and the workers, and returns when
// all are done.
class WorkPool {
public:
    explicit WorkPool(int threads) {
        for (int t = 1; t < threads; ++t) workers_.emplace_back([this] { loop(); });
    }
    ~WorkPool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    size_t size() const { return workers_.size() + 1; }
    void run(size_t n, const std::function<void(size_t)>& f) {
        if (workers_.empty() || n < 2) {
            for (size_t i = 0; i < n; ++i) f(i);
            return;
        }
        {
            std::lock_guard<std::mutex> g(m_);
            job_ = &f;
            n_ = n;
            next_.store(0);
            busy_ = workers_.size();
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> g(m_);
        done_cv_.wait(g, [this] { return busy_ == 0; });
        job_ = nullptr;
    }

private:
    void work() {
        for (size_t i = next_.fetch_add(1); i < n_; i = next_.fetch_add(1)) (*job_)(i);
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            work();
            std::lock_guard<std::mutex> g(m_);
            if (--busy_ == 0) done_cv_.notify_one();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(size_t)>* job_ = nullptr;
    size_t n_ = 0;
    std::atomic<size_t> next_{0};
    size_t busy_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

struct Result {
    uint64_t warmup_proposed = 0, warmup_accepted = 0, production_proposed = 0, production_accepted = 0;
    std::vector<uint64_t> warmup_chosen, production_chosen;
    Stats warmup_retained, production_retained;
    std::vector<Stats> order_signal;
};

struct StepWork {  // MCMCMCStepWorkspace (mcmcmc_sampler.h:114-122)
    std::vector<double> fluxes, weights, order_signal, cond, dens, comp;
};

struct Chain {
    cdet_mcmcmc_config cfg{};        // this chain's settings (its envelope order)
    const MaskPlan* plan = nullptr;  // mask plan of that envelope order
    StepWork work;
    std::string error;
    size_t index = 0;
    uint64_t seed = 0;
    std::mt19937_64 rng;
    std::vector<cdet_vertex> vertices, next;
    std::vector<double> lambdas;
    std::vector<uint64_t> warmup_block_counts;
    Result result;
    bool entered_production = false;
    uint64_t next_checkpoint = std::numeric_limits<uint64_t>::max();
    uint64_t steps_this_run = 0;
    bool stopped = false;
};

void init_result(Result& r, const cdet_mcmcmc_config& c) {
    r = Result{};
    r.warmup_chosen.assign(static_cast<size_t>(c.envelope_order + 1), 0);
    r.production_chosen.assign(static_cast<size_t>(c.envelope_order + 1), 0);
    r.order_signal.assign(static_cast<size_t>(c.envelope_order + 1), {});
}

std::filesystem::path checkpoint_path(const cdet_mcmcmc_config& c, size_t index) {
    std::ostringstream name;
    name << "chain_" << std::setw(6) << std::setfill('0') << index << ".chk";
    return std::filesystem::path(c.checkpoint_dir) / name.str();
}

// ---- checkpoint V1 text format (mcmcmc_sampler.h:718-866) ----------------
void write_stats(std::ostream& out, const Stats& s) {
    out << s.count << ' ' << std::hexfloat << s.mean << ' ' << s.m2 << std::defaultfloat << '\n';
}
void write_u64s(std::ostream& out, const char* label, const std::vector<uint64_t>& v) {
    out << label << ' ' << v.size();
    for (uint64_t x : v) out << ' ' << x;
    out << '\n';
}
void write_doubles(std::ostream& out, const char* label, const std::vector<double>& v) {
    out << label << ' ' << v.size() << std::hexfloat;
    for (double x : v) out << ' ' << x;
    out << std::defaultfloat << '\n';
}

void save_checkpoint(const cdet_mcmcmc_config& c, const Chain& ch) {
    if (!c.checkpoint_enabled) return;
    std::filesystem::create_directories(c.checkpoint_dir);
    const std::filesystem::path path = checkpoint_path(c, ch.index);
    const std::filesystem::path tmp = path.string() + ".tmp";
    std::ofstream out(tmp, std::ios::out | std::ios::trunc);
    if (!out.is_open()) throw std::runtime_error("Could not open MCMCMC checkpoint for writing: " + tmp.string());
    out << std::setprecision(std::numeric_limits<long double>::max_digits10);
    out << "CDET_MCMCMC_CHAIN_CHECKPOINT_V1\n";
    out << "chain_index " << ch.index << '\n';
    out << "seed " << ch.seed << '\n';
    out << "order_min " << c.order_min << '\n';
    out << "order_max " << c.order_max << '\n';
    out << "envelope_order " << c.envelope_order << '\n';
    out << "normalization_order " << c.normalization_order << '\n';
    out << "lambda_warmup_steps " << c.lambda_warmup_steps << '\n';
    out << "production_steps " << c.production_steps << '\n';
    out << "seed_mode " << seed_mode_name(c.seed_mode) << '\n';
    out << "rng_state\n" << ch.rng << '\n';
    write_doubles(out, "lambdas", ch.lambdas);
    write_u64s(out, "warmup_block_counts", ch.warmup_block_counts);
    out << "vertices " << ch.vertices.size() << '\n' << std::hexfloat;
    for (const auto& v : ch.vertices) out << v.tau << ' ' << v.x << ' ' << v.y << ' ' << v.replica << '\n';
    out << std::defaultfloat;
    const Result& r = ch.result;
    out << "warmup_proposed_steps " << r.warmup_proposed << '\n';
    out << "warmup_accepted_steps " << r.warmup_accepted << '\n';
    out << "production_proposed_steps " << r.production_proposed << '\n';
    out << "production_accepted_steps " << r.production_accepted << '\n';
    write_u64s(out, "warmup_chosen_counts", r.warmup_chosen);
    write_u64s(out, "production_chosen_counts", r.production_chosen);
    out << "warmup_retained_fraction_stats ";
    write_stats(out, r.warmup_retained);
    out << "production_retained_fraction_stats ";
    write_stats(out, r.production_retained);
    out << "order_signal_stats " << r.order_signal.size() << '\n';
    for (const auto& s : r.order_signal) write_stats(out, s);
    out << "END\n";
    out.close();
    if (!out) throw std::runtime_error("Failed while writing MCMCMC checkpoint: " + tmp.string());
    std::filesystem::rename(tmp, path);
}

long double parse_float_token(const std::string& tok, const char* ctx) {
    errno = 0;
    char* end = nullptr;
    const long double v = std::strtold(tok.c_str(), &end);
    if (end == tok.c_str() || *end != '\0' || errno == ERANGE)
        throw std::runtime_error("Could not parse MCMCMC checkpoint floating value for " + std::string(ctx) + ": " + tok);
    return v;
}
long double read_float(std::istream& in, const char* ctx) {
    std::string tok;
    in >> tok;
    if (!in) throw std::runtime_error("Missing MCMCMC checkpoint floating value for " + std::string(ctx));
    return parse_float_token(tok, ctx);
}
void expect_label(std::istream& in, const char* label) {
    std::string l;
    in >> l;
    if (l != label)
        throw std::runtime_error("MCMCMC checkpoint expected label " + std::string(label) + " but found " + l);
}
template <typename T>
T read_scalar(std::istream& in, const char* label) {
    expect_label(in, label);
    T v{};
    in >> v;
    return v;
}
void read_stats(std::istream& in, Stats& s) {
    in >> s.count;
    if (!in) throw std::runtime_error("Malformed MCMCMC checkpoint running-stat count");
    s.mean = read_float(in, "running-stat mean");
    s.m2 = read_float(in, "running-stat m2");
}
void read_u64s(std::istream& in, const char* label, std::vector<uint64_t>& v) {
    std::string l;
    size_t n = 0;
    in >> l >> n;
    if (l != label) throw std::runtime_error("MCMCMC checkpoint expected label " + std::string(label) + " but found " + l);
    v.assign(n, 0);
    for (auto& x : v) in >> x;
}
void read_doubles(std::istream& in, const char* label, std::vector<double>& v) {
    std::string l;
    size_t n = 0;
    in >> l >> n;
    if (l != label) throw std::runtime_error("MCMCMC checkpoint expected label " + std::string(label) + " but found " + l);
    v.assign(n, 0.0);
    for (auto& x : v) x = static_cast<double>(read_float(in, label));
}

bool load_checkpoint(const cdet_mcmcmc_config& c, size_t index, uint64_t seed, Chain& ch) {  // :868-982
    const std::filesystem::path path = checkpoint_path(c, index);
    std::ifstream in(path);
    if (!in.is_open()) return false;
    std::string magic;
    in >> magic;
    if (magic != "CDET_MCMCMC_CHAIN_CHECKPOINT_V1")
        throw std::runtime_error("Unsupported MCMCMC checkpoint format: " + path.string());
    ch.index = read_scalar<size_t>(in, "chain_index");
    ch.seed = read_scalar<uint64_t>(in, "seed");
    const int omin = read_scalar<int>(in, "order_min");
    const int omax = read_scalar<int>(in, "order_max");
    const int env = read_scalar<int>(in, "envelope_order");
    const int norm = read_scalar<int>(in, "normalization_order");
    const int warm = read_scalar<int>(in, "lambda_warmup_steps");
    const int prod = read_scalar<int>(in, "production_steps");
    expect_label(in, "seed_mode");
    std::string smode;
    in >> smode;
    if (ch.index != index || ch.seed != seed || omin != c.order_min || omax != c.order_max ||
        env != c.envelope_order || norm != c.normalization_order || warm != c.lambda_warmup_steps ||
        prod > c.production_steps || smode != seed_mode_name(c.seed_mode))
        throw std::runtime_error("MCMCMC checkpoint does not match current run settings: " + path.string());
    expect_label(in, "rng_state");
    in >> ch.rng;
    read_doubles(in, "lambdas", ch.lambdas);
    read_u64s(in, "warmup_block_counts", ch.warmup_block_counts);
    expect_label(in, "vertices");
    size_t nv = 0;
    in >> nv;
    ch.vertices.assign(nv, cdet_vertex{});
    for (auto& v : ch.vertices) {
        v.tau = static_cast<double>(read_float(in, "vertex tau"));
        v.x = static_cast<double>(read_float(in, "vertex x"));
        v.y = static_cast<double>(read_float(in, "vertex y"));
        v.replica = static_cast<double>(read_float(in, "vertex replica"));
    }
    init_result(ch.result, c);
    Result& r = ch.result;
    r.warmup_proposed = read_scalar<uint64_t>(in, "warmup_proposed_steps");
    r.warmup_accepted = read_scalar<uint64_t>(in, "warmup_accepted_steps");
    r.production_proposed = read_scalar<uint64_t>(in, "production_proposed_steps");
    r.production_accepted = read_scalar<uint64_t>(in, "production_accepted_steps");
    read_u64s(in, "warmup_chosen_counts", r.warmup_chosen);
    read_u64s(in, "production_chosen_counts", r.production_chosen);
    expect_label(in, "warmup_retained_fraction_stats");
    read_stats(in, r.warmup_retained);
    expect_label(in, "production_retained_fraction_stats");
    read_stats(in, r.production_retained);
    expect_label(in, "order_signal_stats");
    size_t ns = 0;
    in >> ns;
    r.order_signal.assign(ns, {});
    for (auto& s : r.order_signal) read_stats(in, s);
    std::string end;
    in >> end;
    if (end != "END") throw std::runtime_error("Malformed MCMCMC checkpoint footer: " + path.string());
    return true;
}

void adapt_lambdas(std::vector<double>& lambdas, const cdet_mcmcmc_config& c, const std::vector<uint64_t>& counts) {
    uint64_t total = 0;
    for (int o = c.order_min; o <= c.order_max; ++o) total += counts[static_cast<size_t>(o)];
    if (total == 0) return;
    const int oc = c.order_max - c.order_min + 1;
    const double target = 1.0 / static_cast<double>(oc);
    for (int o = c.order_min; o <= c.order_max; ++o) {
        const double actual = (static_cast<double>(counts[static_cast<size_t>(o)]) + 1.0) /
                              (static_cast<double>(total) + static_cast<double>(oc));
        const double update = std::exp(c.lambda_adapt_rate * std::log(target / actual));
        lambdas[static_cast<size_t>(o)] *= update;
    }
}

uint64_t completed_steps(const Chain& ch) { return ch.result.warmup_proposed + ch.result.production_proposed; }

void checkpoint_if_due(const cdet_mcmcmc_config& c, Chain& ch, bool force) {
    if (!c.checkpoint_enabled) return;
    const uint64_t done = completed_steps(ch);
    if (force || done >= ch.next_checkpoint) {
        save_checkpoint(c, ch);
        if (!force)
            while (ch.next_checkpoint <= done) ch.next_checkpoint += c.checkpoint_interval_steps;
    }
}

// One heat-bath step on precomputed integrand values (run_mcmcmc_step, :524-582).
struct StepOut {
    uint64_t chosen_mask;
    int chosen_order;
};

StepOut heat_bath_step(const cdet_mcmcmc_config& c, const MaskPlan& plan, const SeedKernel& seed, Chain& ch,
                       const double* values, const double* cond_values) {
    StepWork& ws = ch.work;
    std::vector<double>& fluxes = ws.fluxes;
    std::vector<double>& weights = ws.weights;
    std::vector<double>& order_signal = ws.order_signal;
    const size_t nc = plan.masks.size();
    fluxes.resize(nc);
    weights.resize(nc);
    order_signal.assign(static_cast<size_t>(c.envelope_order + 1), 0.0);
    // conditional generation probabilities: given (the batch callback) or
    // computed here
    const double* condp = cond_values;
    if (!seed.uniform && !condp) {
        conditional_probabilities(seed, ch.vertices.data(), c.envelope_order, ws.cond, ws.dens, ws.comp);
        condp = ws.cond.data();
    }
    double norm = 0.0;
    for (size_t k = 0; k < nc; ++k) {
        const uint64_t mask = plan.masks[k];
        const int order = plan.orders[k];
        const double value = values[mask];
        // Probability of regenerating the dropped vertices from this subset.
        const double cond = seed.uniform ? 1.0 : condp[plan.full_mask ^ mask];
        const double flux = ch.lambdas[static_cast<size_t>(order)] * value * cond;
        const double w = std::abs(flux) * plan.inv_binom[static_cast<size_t>(order)];
        fluxes[k] = flux;
        weights[k] = w;
        norm += w;
    }
    // sample_weighted_candidate_mask (:222-248)
    if (!(norm > 0.0) || !std::isfinite(norm)) throw std::runtime_error("MCMCMC local normalization is zero or non-finite");
    std::uniform_real_distribution<double> uniform(0.0, norm);
    const double threshold = uniform(ch.rng);
    double cumulative = 0.0;
    uint64_t last_positive = 0;
    uint64_t chosen = 0;
    bool found = false;
    // The order signals (:563-571) accumulate in the same pass as the
    // cumulative search, each sum in mask order as in two separate loops: the
    // search's serial chain of additions then overlaps the signal work.
    size_t k = 0;
    for (; k < nc && !found; ++k) {
        const int order = plan.orders[k];
        order_signal[static_cast<size_t>(order)] += (fluxes[k] / norm) * plan.inv_binom[static_cast<size_t>(order)];
        const double w = weights[k];
        if (!(w > 0.0)) continue;
        if (std::isfinite(w)) last_positive = plan.masks[k];
        cumulative += w;
        if (cumulative > threshold) {
            chosen = plan.masks[k];
            found = true;
        }
    }
    for (; k < nc; ++k) {
        const int order = plan.orders[k];
        order_signal[static_cast<size_t>(order)] += (fluxes[k] / norm) * plan.inv_binom[static_cast<size_t>(order)];
    }
    if (!found) chosen = last_positive;
    const int chosen_order = __builtin_popcountll(chosen);
    // create_next_vertices_from_seed_into (:380-399): kept vertices in index
    // order, then each regenerated vertex is a child of a uniformly drawn
    // vertex of the list so far.
    ch.next.clear();
    for (int i = 0; i < c.envelope_order; ++i)
        if (chosen & (uint64_t{1} << i)) ch.next.push_back(ch.vertices[static_cast<size_t>(i)]);
    while (static_cast<int>(ch.next.size()) < c.envelope_order) {
        std::uniform_int_distribution<int> parent(0, static_cast<int>(ch.next.size()) - 1);
        const size_t pi = static_cast<size_t>(parent(ch.rng));
        const cdet_vertex child = seed.child(ch.next[pi], ch.rng);
        ch.next.push_back(child);
    }
    return StepOut{chosen, chosen_order};
}

int host_thread_count(const cdet_mcmcmc_config& c) {
    if (c.host_threads > 0) return c.host_threads;
    const unsigned hw = std::thread::hardware_concurrency();
    return static_cast<int>(std::max(1U, std::min(16U, hw)));
}

// run_mcmcmc_chains_parallel (mcmcmc_sampler.h:1090-1145) with all live chains
// advanced together: one batched integrand call per step, then the per-chain
// step work (heat-bath choice, regeneration, statistics, checkpoint) spread
// over host threads.  A chain's work touches only its own state, files and
// trace row, so the thread schedule cannot change any result.
//
// envelopes (may be null): chain k's envelope order.  Its sampler then runs
// with envelope_order = envelopes[k] and order_max = min(order_max,
// envelopes[k] - 1), and its checkpoint records those values.  Without it every
// chain uses the configuration as given, as the reference does.
void run_chains(const cdet_mcmcmc_config& c, const SeedKernel& seed, const cdet_histogram_seed* hseed,
                const int32_t* envelopes, const cdet_mcmcmc_backend& be, cdet_mcmcmc_chain_result* results,
                double* ratios, uint64_t* trace, int64_t trace_stride) {
    if (!be.eval) throw std::runtime_error("cdet_mcmcmc_run: null evaluation callback");
    const bool gpu_cond = !seed.uniform && be.cond != nullptr;
    std::vector<Chain> chains(static_cast<size_t>(c.chain_count));
    std::map<int, MaskPlan> plans;  // by envelope order
    int max_env = 0;
    for (int k = 0; k < c.chain_count; ++k) {
        Chain& ch = chains[static_cast<size_t>(k)];
        ch.cfg = c;
        if (envelopes) {
            const int e = envelopes[k];
            ch.cfg.envelope_order = e;
            ch.cfg.order_max = std::min(c.order_max, e - 1);
        }
        validate(ch.cfg);
        if (ch.cfg.envelope_order <= 1)
            throw std::runtime_error("The all-subset replica integrand is defined only for envelope order >= 2");
        if (!plans.count(ch.cfg.envelope_order)) plans.emplace(ch.cfg.envelope_order, make_plan(ch.cfg));
        ch.plan = &plans.at(ch.cfg.envelope_order);
        max_env = std::max(max_env, ch.cfg.envelope_order);
    }
    constexpr uint64_t kChainTag = 0x4D434D434D435348ULL;  // :1103
    for (int k = 0; k < c.chain_count; ++k) {
        Chain& ch = chains[static_cast<size_t>(k)];
        const cdet_mcmcmc_config& cc = ch.cfg;
        const int E = cc.envelope_order;
        const uint64_t s = derive_seed(c.base_seed, kChainTag, static_cast<uint64_t>(k));
        bool resumed = false;
        if (c.checkpoint_enabled && c.resume_mode != CDET_RESUME_OFF) {
            resumed = load_checkpoint(cc, static_cast<size_t>(k), s, ch);
            if (!resumed && c.resume_mode == CDET_RESUME_REQUIRE)
                throw std::runtime_error("MCMCMC_RESUME=require but checkpoint is missing for chain " +
                                         std::to_string(k));
        }
        if (!resumed) {  // make_initial_mcmcmc_chain_state (:696-709)
            ch.index = static_cast<size_t>(k);
            ch.seed = s;
            ch.rng.seed(s);
            ch.vertices.clear();
            for (int i = 0; i < E; ++i) ch.vertices.push_back(uniform_vertex(ch.rng));
            ch.lambdas.assign(c.lambdas, c.lambdas + E + 1);
            ch.warmup_block_counts.assign(static_cast<size_t>(E + 1), 0);
            init_result(ch.result, cc);
        }
        if (c.checkpoint_enabled) {
            const uint64_t done = completed_steps(ch);
            ch.next_checkpoint = ((done / c.checkpoint_interval_steps) + 1U) * c.checkpoint_interval_steps;
        }
    }
    if (c.checkpoint_enabled) std::filesystem::create_directories(c.checkpoint_dir);
    auto enter_production = [&](Chain& ch) {
        if (ch.entered_production) return;
        if (c.lambda_warmup_steps % c.lambda_adapt_interval != 0 &&
            std::any_of(ch.warmup_block_counts.begin(), ch.warmup_block_counts.end(),
                        [](uint64_t x) { return x != 0; })) {
            adapt_lambdas(ch.lambdas, ch.cfg, ch.warmup_block_counts);
            std::fill(ch.warmup_block_counts.begin(), ch.warmup_block_counts.end(), 0);
        }
        ch.entered_production = true;
    };
    auto live = [&](const Chain& ch) {
        if (ch.stopped) return false;
        return ch.result.warmup_proposed < static_cast<uint64_t>(c.lambda_warmup_steps) ||
               ch.result.production_proposed < static_cast<uint64_t>(c.production_steps);
    };
    // One chain's step on its integrand values (:1030-1085).
    auto advance = [&](Chain& ch, const double* values, const double* cond_values) {
        const cdet_mcmcmc_config& cc = ch.cfg;
        const int E = cc.envelope_order;
        const bool warm = ch.result.warmup_proposed < static_cast<uint64_t>(c.lambda_warmup_steps);
        const StepOut s = heat_bath_step(cc, *ch.plan, seed, ch, values, cond_values);
        ch.vertices.swap(ch.next);
        if (trace && static_cast<int64_t>(ch.steps_this_run) < trace_stride)
            trace[static_cast<int64_t>(ch.index) * trace_stride + static_cast<int64_t>(ch.steps_this_run)] = s.chosen_mask;
        ++ch.steps_this_run;
        Result& r = ch.result;
        const std::vector<double>& order_signal = ch.work.order_signal;
        if (warm) {  // :1038-1055
            const uint64_t next_step = r.warmup_proposed + 1;
            ++r.warmup_proposed;
            ++r.warmup_accepted;
            ++r.warmup_chosen[static_cast<size_t>(s.chosen_order)];
            r.warmup_retained.add(static_cast<long double>(s.chosen_order) / static_cast<long double>(E));
            ++ch.warmup_block_counts[static_cast<size_t>(s.chosen_order)];
            if (next_step % static_cast<uint64_t>(c.lambda_adapt_interval) == 0) {
                adapt_lambdas(ch.lambdas, cc, ch.warmup_block_counts);
                std::fill(ch.warmup_block_counts.begin(), ch.warmup_block_counts.end(), 0);
            }
        } else {  // :1064-1083
            ++r.production_proposed;
            ++r.production_accepted;
            ++r.production_chosen[static_cast<size_t>(s.chosen_order)];
            r.production_retained.add(static_cast<long double>(s.chosen_order) / static_cast<long double>(E));
            for (int o = cc.order_min; o <= cc.order_max; ++o)
                r.order_signal[static_cast<size_t>(o)].add(order_signal[static_cast<size_t>(o)] /
                                                             ch.lambdas[static_cast<size_t>(o)]);
        }
        checkpoint_if_due(cc, ch, false);
        if (c.stop_after_steps && ch.steps_this_run >= c.stop_after_steps) ch.stopped = true;
    };
    // Threads pay off once a step's host work is a few microseconds per chain:
    // the histogram seed's subset recursion (2^E * E) or a large mask plan.
    size_t host_work = 0;
    for (const Chain& ch : chains)
        host_work += ch.plan->masks.size() + ((seed.uniform || gpu_cond) ? 0
                                                 : (size_t{1} << ch.cfg.envelope_order) *
                                                       static_cast<size_t>(ch.cfg.envelope_order));
    const int threads = host_work >= 32768 ? std::min(host_thread_count(c), c.chain_count) : 1;
    WorkPool pool(threads);
    // Sweeps.  With an asynchronous backend (submit + wait) the chains form
    // two groups (the halves of the chain list), each with its own batch slot:
    // while one group's batch is evaluated, the other group's chains take
    // their heat-bath steps and its next batch is submitted.  Every chain
    // takes the same steps as in the lockstep form, since a chain's values
    // do not depend on the batch it is evaluated in.
    const bool async = be.submit != nullptr && be.wait != nullptr && c.chain_count >= 2;
    const int ngroups = async ? 2 : 1;
    const size_t half = static_cast<size_t>((c.chain_count + 1) / 2);
    // Value buffer of a group: from the backend's allocator when it has one
    // (pinned memory for a GPU backend), else a vector.
    struct Values {
        const cdet_mcmcmc_backend* be = nullptr;
        std::vector<double> vec;
        double* buf = nullptr;
        size_t cap = 0;
        Values() = default;
        Values(const Values&) = delete;
        Values& operator=(const Values&) = delete;
        ~Values() {
            if (buf) be->host_free(be->eval_user, buf);
        }
        // room for n values; every value is written by the evaluation callback
        double* ensure(size_t n) {
            if (!be || !be->host_alloc || !be->host_free) {
                vec.resize(n);
                return vec.data();
            }
            if (n > cap) {
                if (buf) be->host_free(be->eval_user, buf);
                buf = nullptr;
                cap = 0;
                buf = static_cast<double*>(be->host_alloc(be->eval_user, sizeof(double) * std::max<size_t>(n, 1)));
                if (!buf) throw std::runtime_error("integrand value buffer allocation failed");
                cap = n;
            }
            return buf;
        }
    };
    struct Group {
        std::vector<cdet_vertex> hv;
        std::vector<int32_t> ho;
        Values hout;
        double* out = nullptr;  // hout's values of the current batch
        std::vector<double> hcond;
        std::vector<size_t> batch, hoff;
        std::vector<size_t> by_size;  // positions in batch, largest envelope first
        size_t lo = 0, hi = 0;  // chain index range
        bool in_flight = false;
    } groups[2];
    for (int g = 0; g < ngroups; ++g) {
        groups[g].hout.be = &be;
        groups[g].lo = ngroups == 1 ? 0 : (g == 0 ? 0 : half);
        groups[g].hi = ngroups == 1 ? chains.size() : (g == 0 ? half : chains.size());
    }
    // the group's live chains and their envelopes; false when none is left
    auto prepare = [&](Group& gr) {
        gr.batch.clear();
        for (size_t k = gr.lo; k < gr.hi; ++k) {
            Chain& ch = chains[k];
            if (!live(ch)) continue;
            if (ch.result.warmup_proposed >= static_cast<uint64_t>(c.lambda_warmup_steps)) enter_production(ch);
            gr.batch.push_back(ch.index);
        }
        if (gr.batch.empty()) return false;
        gr.hv.clear();
        gr.ho.clear();
        gr.hoff.assign(1, 0);
        for (size_t b : gr.batch) {
            const Chain& ch = chains[b];
            gr.hv.insert(gr.hv.end(), ch.vertices.begin(), ch.vertices.end());
            gr.ho.push_back(ch.cfg.envelope_order);
            gr.hoff.push_back(gr.hoff.back() + (size_t{1} << ch.cfg.envelope_order));
        }
        gr.out = gr.hout.ensure(gr.hoff.back());
        // the host threads take the chains largest first, so that the last
        // ones to start are short (a stable counting sort by envelope order)
        gr.by_size.resize(gr.batch.size());
        {
            size_t pos[kMaxEnvelope + 2] = {};
            for (int e : gr.ho) ++pos[kMaxEnvelope - e + 1];
            for (int e = 1; e <= kMaxEnvelope + 1; ++e) pos[e] += pos[e - 1];
            for (size_t bi = 0; bi < gr.ho.size(); ++bi) gr.by_size[pos[kMaxEnvelope - gr.ho[bi]]++] = bi;
        }
        return true;
    };
    auto submit = [&](int g) {
        Group& gr = groups[g];
        const int rc = be.submit(be.eval_user, g, gr.hv.data(), gr.ho.data(), static_cast<int64_t>(gr.batch.size()));
        if (rc != 0) throw std::runtime_error("integrand evaluation failed (status " + std::to_string(rc) + ")");
        gr.in_flight = true;
    };
    // values of the group's batch (waiting for it, or evaluating it now), then
    // the chains' steps
    auto complete = [&](int g) {
        Group& gr = groups[g];
        int rc = 0;
        if (async) {
            gr.in_flight = false;
            rc = be.wait(be.eval_user, g, gr.out);
        } else {
            rc = be.eval(be.eval_user, gr.hv.data(), gr.ho.data(), static_cast<int64_t>(gr.batch.size()),
                         gr.out);
        }
        if (rc != 0) throw std::runtime_error("integrand evaluation failed (status " + std::to_string(rc) + ")");
        if (gpu_cond) {
            gr.hcond.resize(gr.hoff.back());  // every value is written by the callback
            const int rcc = be.cond(be.cond_user, hseed, gr.hv.data(), gr.ho.data(),
                                    static_cast<int64_t>(gr.batch.size()), gr.hcond.data());
            if (rcc != 0)
                throw std::runtime_error("conditional generation probabilities failed (status " + std::to_string(rcc) +
                                         ")");
        }
        pool.run(gr.batch.size(), [&](size_t i) {
            const size_t bi = gr.by_size[i];
            Chain& ch = chains[gr.batch[bi]];
            try {
                advance(ch, gr.out + gr.hoff[bi], gpu_cond ? gr.hcond.data() + gr.hoff[bi] : nullptr);
            } catch (const std::exception& e) {
                ch.error = e.what();
                ch.stopped = true;
            }
        });
        for (size_t b : gr.batch)  // the lowest failing chain reports, as a serial loop would
            if (!chains[b].error.empty()) throw std::runtime_error(chains[b].error);
    };
    // a failure leaves no batch in flight behind
    auto drain = [&]() {
        for (int g = 0; g < ngroups; ++g)
            if (groups[g].in_flight) {
                groups[g].in_flight = false;
                (void)be.wait(be.eval_user, g, groups[g].out);
            }
    };
    try {
        bool active[2] = {false, false};
        for (int g = 0; g < ngroups; ++g) {
            active[g] = prepare(groups[g]);
            if (active[g] && async) submit(g);
        }
        while (active[0] || active[1]) {
            for (int g = 0; g < ngroups; ++g) {
                if (!active[g]) continue;
                complete(g);
                active[g] = prepare(groups[g]);
                if (active[g] && async) submit(g);
            }
        }
    } catch (...) {
        drain();
        throw;
    }
    for (auto& ch : chains) {
        if (!ch.stopped) {
            if (ch.result.warmup_proposed >= static_cast<uint64_t>(c.lambda_warmup_steps)) enter_production(ch);
            checkpoint_if_due(ch.cfg, ch, true);  // forced final save (:1085)
        }
    }
    if (ratios) {
        for (int o = 0; o <= max_env; ++o) ratios[o] = std::numeric_limits<double>::quiet_NaN();
        const size_t norm = static_cast<size_t>(c.normalization_order);
        if (!envelopes) {
            // Cross-chain statistics (:1117-1142): ratio of the cross-chain
            // means of the order signals.
            std::vector<Stats> cross(static_cast<size_t>(max_env + 1));
            for (auto& ch : chains)
                for (int o = c.order_min; o <= c.order_max; ++o) {
                    const Stats& s = ch.result.order_signal[static_cast<size_t>(o)];
                    if (s.count > 0) cross[static_cast<size_t>(o)].add(s.mean);
                }
            const long double ref = cross[norm].mean;
            for (int o = c.order_min; o <= c.order_max; ++o)
                ratios[o] = ref == 0.0L ? std::numeric_limits<double>::quiet_NaN()
                                        : static_cast<double>(cross[static_cast<size_t>(o)].mean / ref);
        } else {
            // Chains of different envelopes normalise their signals by
            // different partition sums, so each chain's ratio to its own
            // normalization-order signal is formed first and those ratios
            // are averaged over the chains that measure the order.
            std::vector<Stats> cross(static_cast<size_t>(max_env + 1));
            for (auto& ch : chains) {
                const long double ref = ch.result.order_signal[norm].mean;
                if (ch.result.order_signal[norm].count == 0 || ref == 0.0L) continue;
                for (int o = ch.cfg.order_min; o <= ch.cfg.order_max; ++o) {
                    const Stats& s = ch.result.order_signal[static_cast<size_t>(o)];
                    if (s.count > 0) cross[static_cast<size_t>(o)].add(s.mean / ref);
                }
            }
            for (int o = c.order_min; o <= std::min(c.order_max, max_env - 1); ++o)
                if (cross[static_cast<size_t>(o)].count > 0)
                    ratios[o] = static_cast<double>(cross[static_cast<size_t>(o)].mean);
        }
    }
    if (results) {
        for (auto& ch : chains) {
            const int E = ch.cfg.envelope_order;
            cdet_mcmcmc_chain_result& o = results[ch.index];
            std::memset(&o, 0, sizeof(o));
            o.seed = ch.seed;
            o.warmup_proposed_steps = ch.result.warmup_proposed;
            o.production_proposed_steps = ch.result.production_proposed;
            for (int k = 0; k <= E; ++k) {
                o.warmup_chosen_counts[k] = ch.result.warmup_chosen[static_cast<size_t>(k)];
                o.production_chosen_counts[k] = ch.result.production_chosen[static_cast<size_t>(k)];
                o.final_lambdas[k] = ch.lambdas[static_cast<size_t>(k)];
                const Stats& s = ch.result.order_signal[static_cast<size_t>(k)];
                o.order_signal_mean[k] = static_cast<double>(s.mean);
                o.order_signal_stderr[k] = static_cast<double>(s.stderr_());
                o.order_signal_count[k] = s.count;
            }
            o.retained_fraction_mean = static_cast<double>(ch.result.production_retained.mean);
        }
    }
}

}  // namespace

extern "C" {

const char* cdet_mcmcmc_last_error(void) { return g_error.c_str(); }

void cdet_mcmcmc_config_defaults(cdet_mcmcmc_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->order_min = 2;
    c->order_max = 4;
    c->envelope_order = 5;
    c->normalization_order = 2;
    for (int o = 0; o <= CDET_MCMCMC_MAX_ENVELOPE; ++o) c->lambdas[o] = 1.0;
    c->lambda_warmup_steps = 0;
    c->lambda_adapt_interval = 1000;
    c->lambda_adapt_rate = 0.25;
    c->production_steps = 100000;
    c->chain_count = 1;
    c->base_seed = 1;
    c->seed_mode = CDET_SEED_UNIFORM;
    c->checkpoint_enabled = 0;
    c->checkpoint_interval_steps = 10000000ULL;
    c->resume_mode = CDET_RESUME_OFF;
    std::strncpy(c->checkpoint_dir, "mcmcmc_checkpoints", sizeof(c->checkpoint_dir) - 1);
    c->N1 = 1;
    c->N2 = 1;
    c->beta = 1.0;
    c->histogram_learning_steps = 0;
    c->histogram_tau_bins = 16;
    c->histogram_uniform_mix = 0.05;
    c->histogram_pseudo_count = 0.0;
}

int cdet_mcmcmc_run_seeded(const cdet_mcmcmc_config* cfg, const cdet_histogram_seed* seed, cdet_batch_eval_fn eval,
                           void* eval_user, cdet_mcmcmc_chain_result* results, double* ratios, uint64_t* trace,
                           int64_t trace_stride) {
    try {
        if (!cfg || !eval) throw std::runtime_error("cdet_mcmcmc_run: null argument");
        validate(*cfg);
        const SeedKernel kernel = SeedKernel::from(seed);
        run_chains(*cfg, kernel, seed, nullptr, cdet_mcmcmc_backend{eval, eval_user, nullptr, nullptr}, results,
                   ratios, trace, trace_stride);
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int cdet_mcmcmc_run_envelopes(const cdet_mcmcmc_config* cfg, const cdet_histogram_seed* seed,
                              const int32_t* envelope_orders, cdet_batch_eval_fn eval, void* eval_user,
                              cdet_mcmcmc_chain_result* results, double* ratios, uint64_t* trace,
                              int64_t trace_stride) {
    try {
        if (!cfg || !eval || !envelope_orders) throw std::runtime_error("cdet_mcmcmc_run_envelopes: null argument");
        validate(*cfg);
        const SeedKernel kernel = SeedKernel::from(seed);
        run_chains(*cfg, kernel, seed, envelope_orders, cdet_mcmcmc_backend{eval, eval_user, nullptr, nullptr},
                   results, ratios, trace, trace_stride);
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int cdet_mcmcmc_run_backend(const cdet_mcmcmc_config* cfg, const cdet_histogram_seed* seed,
                            const int32_t* envelope_orders, const cdet_mcmcmc_backend* backend,
                            cdet_mcmcmc_chain_result* results, double* ratios, uint64_t* trace,
                            int64_t trace_stride) {
    try {
        if (!cfg || !backend) throw std::runtime_error("cdet_mcmcmc_run_backend: null argument");
        validate(*cfg);
        const SeedKernel kernel = SeedKernel::from(seed);
        run_chains(*cfg, kernel, seed, envelope_orders, *backend, results, ratios, trace, trace_stride);
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int cdet_histogram_learn(const cdet_mcmcmc_config* cfg, cdet_batch_eval_fn eval, void* eval_user,
                         cdet_histogram_seed* seed, double* total_pair_weight) {
    try {
        if (!cfg || !eval || !seed || !seed->tau_prob || !seed->space_prob)
            throw std::runtime_error("cdet_histogram_learn: null argument");
        if (cfg->envelope_order < 2 || cfg->envelope_order > CDET_MCMCMC_MAX_ENVELOPE)
            throw std::runtime_error("cdet_histogram_learn: envelope order out of range");
        learn_histogram(*cfg, eval, eval_user, *seed, total_pair_weight);
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int cdet_histogram_export(const cdet_histogram_seed* seed, double beta, const char* prefix) {
    try {
        if (!seed || !seed->tau_prob || !seed->space_prob) throw std::runtime_error("cdet_histogram_export: null argument");
        export_histogram(*seed, beta, prefix ? std::string(prefix) : std::string());
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int cdet_mcmcmc_run(const cdet_mcmcmc_config* cfg, cdet_batch_eval_fn eval, void* eval_user,
                    cdet_mcmcmc_chain_result* results, double* ratios, uint64_t* trace, int64_t trace_stride) {
    try {
        if (!cfg || !eval) throw std::runtime_error("cdet_mcmcmc_run: null argument");
        const cdet_mcmcmc_config& c = *cfg;
        validate(c);
        if (c.seed_mode != CDET_SEED_HISTOGRAM) {
            run_chains(c, SeedKernel{}, nullptr, nullptr, cdet_mcmcmc_backend{eval, eval_user, nullptr, nullptr},
                       results, ratios, trace, trace_stride);
            return 0;
        }
        // src/main.cpp:1071-1102: learn, export, optionally stop, run.
        std::vector<double> tau(static_cast<size_t>(c.histogram_tau_bins) * static_cast<size_t>(c.histogram_tau_bins));
        const size_t sites = static_cast<size_t>(std::max(c.N1, 0)) * static_cast<size_t>(std::max(c.N2, 0));
        std::vector<double> space(sites * sites);
        cdet_histogram_seed h{};
        h.tau_prob = tau.data();
        h.space_prob = space.data();
        learn_histogram(c, eval, eval_user, h, nullptr);
        export_histogram(h, c.beta, std::string(c.histogram_output_prefix));
        if (c.histogram_learning_only) return 0;
        run_chains(c, SeedKernel::from(&h), &h, nullptr, cdet_mcmcmc_backend{eval, eval_user, nullptr, nullptr},
                   results, ratios, trace, trace_stride);
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int cdet_seed_conditional_probabilities(const cdet_histogram_seed* seed, const cdet_vertex* v, int E, double* out) {
    try {
        if (!v || !out) throw std::runtime_error("cdet_seed_conditional_probabilities: null argument");
        if (E < 1 || E > 30) throw std::runtime_error("cdet_seed_conditional_probabilities: envelope order out of range");
        const SeedKernel k = SeedKernel::from(seed);
        std::vector<double> p, dens, comp;
        conditional_probabilities(k, v, E, p, dens, comp);
        std::copy(p.begin(), p.end(), out);
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int cdet_seed_sample_children(const cdet_histogram_seed* seed, const cdet_vertex* parents, int64_t n,
                              uint64_t rng_seed, cdet_vertex* out) {
    try {
        if ((!parents || !out) && n > 0) throw std::runtime_error("cdet_seed_sample_children: null argument");
        const SeedKernel k = SeedKernel::from(seed);
        std::mt19937_64 rng(rng_seed);
        for (int64_t i = 0; i < n; ++i) out[i] = k.child(parents[i], rng);
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int cdet_seed_density(const cdet_histogram_seed* seed, const cdet_vertex* child, const cdet_vertex* parent,
                      double* out) {
    try {
        if (!child || !parent || !out) throw std::runtime_error("cdet_seed_density: null argument");
        *out = SeedKernel::from(seed).density(*child, *parent);
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

}  // extern "C"

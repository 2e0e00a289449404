// cdet_mcmcmc.cpp -- lockstep set heat-bath MCMCMC driver (host).
//
// Reproduces the reference sampler chain by chain
// (include/cdet/mcmcmc_sampler.h:159-1145, include/cdet/mcmcmc_seed.h:41-61):
// every chain keeps its own std::mt19937_64 and consumes it in the
// reference's order (envelope draws, uniform_real_distribution(0, W) for the
// subset choice, the parent draw of the seed regeneration, four uniforms per
// regenerated vertex), so its trajectory and checkpoint bytes equal those of
// the reference's per-chain thread.  What changes is the schedule: all live
// chains advance one step together and their envelopes are evaluated by one
// batched integrand call.
//
// Built with g++/libstdc++ (the reference's standard library) so that the
// distributions and the mt19937_64 stream text are the reference's.

#include "../../include/cdet_mcmcmc.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <limits>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
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
    if (c.envelope_order > 48)
        throw std::runtime_error("MCMCMC_ENVELOPE_ORDER must be <= 48 for the current principal-minor backend");
    if (c.normalization_order < c.order_min || c.normalization_order > c.order_max)
        throw std::runtime_error("MCMCMC_NORMALIZATION_ORDER must lie inside the measured order range");
    if (c.chain_count <= 0) throw std::runtime_error("MCMCMC_CHAIN_COUNT must be positive");
    if (c.lambda_warmup_steps < 0 || c.production_steps <= 0)
        throw std::runtime_error("MCMCMC warmup must be nonnegative and production must be positive");
    if (c.lambda_adapt_interval <= 0) throw std::runtime_error("MCMCMC_LAMBDA_ADAPT_INTERVAL must be positive");
    if (c.seed_mode != CDET_SEED_UNIFORM && c.seed_mode != CDET_SEED_HISTOGRAM)
        throw std::runtime_error("MCMCMC_SEED_MODE must be uniform or histogram");
    if (c.seed_mode == CDET_SEED_HISTOGRAM)
        throw std::runtime_error("MCMCMC_SEED_MODE=histogram is not supported by this driver build");
    if (c.resume_mode < CDET_RESUME_OFF || c.resume_mode > CDET_RESUME_REQUIRE)
        throw std::runtime_error("MCMCMC_RESUME must be off, auto, or require");
    if (c.checkpoint_enabled && c.checkpoint_interval_steps == 0)
        throw std::runtime_error("MCMCMC_CHECKPOINT_INTERVAL_STEPS must be positive when checkpointing is enabled");
    for (int o = c.order_min; o <= c.order_max; ++o) {
        const double l = c.lambdas[o];
        if (!(l > 0.0) || !std::isfinite(l)) throw std::runtime_error("MCMCMC_LAMBDAS entries must be positive finite numbers");
    }
}

const char* seed_mode_name(int m) { return m == CDET_SEED_HISTOGRAM ? "histogram" : "uniform"; }

struct Result {
    uint64_t warmup_proposed = 0, warmup_accepted = 0, production_proposed = 0, production_accepted = 0;
    std::vector<uint64_t> warmup_chosen, production_chosen;
    Stats warmup_retained, production_retained;
    std::vector<Stats> order_signal;
};

struct Chain {
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

StepOut heat_bath_step(const cdet_mcmcmc_config& c, const MaskPlan& plan, Chain& ch, const double* values,
                       std::vector<double>& fluxes, std::vector<double>& weights, std::vector<double>& order_signal) {
    const size_t nc = plan.masks.size();
    fluxes.resize(nc);
    weights.resize(nc);
    order_signal.assign(static_cast<size_t>(c.envelope_order + 1), 0.0);
    double norm = 0.0;
    for (size_t k = 0; k < nc; ++k) {
        const uint64_t mask = plan.masks[k];
        const int order = plan.orders[k];
        const double value = values[mask];
        const double cond = 1.0;  // uniform seed: Pgen(full | subset) = 1
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
    for (size_t k = 0; k < nc; ++k) {
        const double w = weights[k];
        if (!(w > 0.0)) continue;
        if (std::isfinite(w)) last_positive = plan.masks[k];
        cumulative += w;
        if (cumulative > threshold) {
            chosen = plan.masks[k];
            found = true;
            break;
        }
    }
    if (!found) chosen = last_positive;
    const int chosen_order = __builtin_popcountll(chosen);
    for (size_t k = 0; k < nc; ++k) {
        const int order = plan.orders[k];
        const double t = fluxes[k] / norm;
        order_signal[static_cast<size_t>(order)] += t * plan.inv_binom[static_cast<size_t>(order)];
    }
    // create_next_vertices_from_seed_into (:380-399) with the uniform seed (mcmcmc_seed.h:41-61)
    ch.next.clear();
    for (int i = 0; i < c.envelope_order; ++i)
        if (chosen & (uint64_t{1} << i)) ch.next.push_back(ch.vertices[static_cast<size_t>(i)]);
    while (static_cast<int>(ch.next.size()) < c.envelope_order) {
        std::uniform_int_distribution<int> parent(0, static_cast<int>(ch.next.size()) - 1);
        (void)parent(ch.rng);
        ch.next.push_back(uniform_vertex(ch.rng));
    }
    return StepOut{chosen, chosen_order};
}

}  // namespace

extern "C" {

const char* cdet_mcmcmc_last_error(void) { return g_error.c_str(); }

int cdet_mcmcmc_run(const cdet_mcmcmc_config* cfg, cdet_batch_eval_fn eval, void* eval_user,
                    cdet_mcmcmc_chain_result* results, double* ratios, uint64_t* trace, int64_t trace_stride) {
    try {
        if (!cfg || !eval) throw std::runtime_error("cdet_mcmcmc_run: null argument");
        const cdet_mcmcmc_config& c = *cfg;
        validate(c);
        if (c.envelope_order <= 1)
            throw std::runtime_error("The all-subset replica integrand is defined only for envelope order >= 2");
        const MaskPlan plan = make_plan(c);
        const int E = c.envelope_order;
        const size_t sc = size_t{1} << E;
        std::vector<Chain> chains(static_cast<size_t>(c.chain_count));
        constexpr uint64_t kChainTag = 0x4D434D434D435348ULL;  // :1103
        for (int k = 0; k < c.chain_count; ++k) {
            Chain& ch = chains[static_cast<size_t>(k)];
            const uint64_t seed = derive_seed(c.base_seed, kChainTag, static_cast<uint64_t>(k));
            bool resumed = false;
            if (c.checkpoint_enabled && c.resume_mode != CDET_RESUME_OFF) {
                resumed = load_checkpoint(c, static_cast<size_t>(k), seed, ch);
                if (!resumed && c.resume_mode == CDET_RESUME_REQUIRE)
                    throw std::runtime_error("MCMCMC_RESUME=require but checkpoint is missing for chain " +
                                             std::to_string(k));
            }
            if (!resumed) {  // make_initial_mcmcmc_chain_state (:696-709)
                ch.index = static_cast<size_t>(k);
                ch.seed = seed;
                ch.rng.seed(seed);
                ch.vertices.clear();
                for (int i = 0; i < E; ++i) ch.vertices.push_back(uniform_vertex(ch.rng));
                ch.lambdas.assign(c.lambdas, c.lambdas + E + 1);
                ch.warmup_block_counts.assign(static_cast<size_t>(E + 1), 0);
                init_result(ch.result, c);
            }
            if (c.checkpoint_enabled) {
                const uint64_t done = completed_steps(ch);
                ch.next_checkpoint = ((done / c.checkpoint_interval_steps) + 1U) * c.checkpoint_interval_steps;
            }
        }
        auto enter_production = [&](Chain& ch) {
            if (ch.entered_production) return;
            if (c.lambda_warmup_steps % c.lambda_adapt_interval != 0 &&
                std::any_of(ch.warmup_block_counts.begin(), ch.warmup_block_counts.end(),
                            [](uint64_t x) { return x != 0; })) {
                adapt_lambdas(ch.lambdas, c, ch.warmup_block_counts);
                std::fill(ch.warmup_block_counts.begin(), ch.warmup_block_counts.end(), 0);
            }
            ch.entered_production = true;
        };
        auto live = [&](const Chain& ch) {
            if (ch.stopped) return false;
            return ch.result.warmup_proposed < static_cast<uint64_t>(c.lambda_warmup_steps) ||
                   ch.result.production_proposed < static_cast<uint64_t>(c.production_steps);
        };
        std::vector<cdet_vertex> hv;
        std::vector<int32_t> ho;
        std::vector<double> hout;
        std::vector<size_t> batch;
        std::vector<double> fluxes, weights, order_signal;
        for (;;) {
            batch.clear();
            for (auto& ch : chains) {
                if (!live(ch)) continue;
                if (ch.result.warmup_proposed >= static_cast<uint64_t>(c.lambda_warmup_steps)) enter_production(ch);
                batch.push_back(ch.index);
            }
            if (batch.empty()) break;
            hv.clear();
            ho.assign(batch.size(), E);
            for (size_t b : batch) hv.insert(hv.end(), chains[b].vertices.begin(), chains[b].vertices.end());
            hout.assign(batch.size() * sc, 0.0);
            const int rc = eval(eval_user, hv.data(), ho.data(), static_cast<int64_t>(batch.size()), hout.data());
            if (rc != 0) throw std::runtime_error("integrand evaluation failed (status " + std::to_string(rc) + ")");
            for (size_t bi = 0; bi < batch.size(); ++bi) {
                Chain& ch = chains[batch[bi]];
                const bool warm = ch.result.warmup_proposed < static_cast<uint64_t>(c.lambda_warmup_steps);
                const StepOut s = heat_bath_step(c, plan, ch, hout.data() + bi * sc, fluxes, weights, order_signal);
                ch.vertices.swap(ch.next);
                if (trace && static_cast<int64_t>(ch.steps_this_run) < trace_stride)
                    trace[static_cast<int64_t>(ch.index) * trace_stride + static_cast<int64_t>(ch.steps_this_run)] =
                        s.chosen_mask;
                ++ch.steps_this_run;
                Result& r = ch.result;
                if (warm) {  // :1038-1055
                    const uint64_t next_step = r.warmup_proposed + 1;
                    ++r.warmup_proposed;
                    ++r.warmup_accepted;
                    ++r.warmup_chosen[static_cast<size_t>(s.chosen_order)];
                    r.warmup_retained.add(static_cast<long double>(s.chosen_order) / static_cast<long double>(E));
                    ++ch.warmup_block_counts[static_cast<size_t>(s.chosen_order)];
                    if (next_step % static_cast<uint64_t>(c.lambda_adapt_interval) == 0) {
                        adapt_lambdas(ch.lambdas, c, ch.warmup_block_counts);
                        std::fill(ch.warmup_block_counts.begin(), ch.warmup_block_counts.end(), 0);
                    }
                } else {  // :1064-1083
                    ++r.production_proposed;
                    ++r.production_accepted;
                    ++r.production_chosen[static_cast<size_t>(s.chosen_order)];
                    r.production_retained.add(static_cast<long double>(s.chosen_order) / static_cast<long double>(E));
                    for (int o = c.order_min; o <= c.order_max; ++o)
                        r.order_signal[static_cast<size_t>(o)].add(order_signal[static_cast<size_t>(o)] /
                                                                     ch.lambdas[static_cast<size_t>(o)]);
                }
                checkpoint_if_due(c, ch, false);
                if (c.stop_after_steps && ch.steps_this_run >= c.stop_after_steps) ch.stopped = true;
            }
        }
        for (auto& ch : chains) {
            if (!ch.stopped) {
                if (ch.result.warmup_proposed >= static_cast<uint64_t>(c.lambda_warmup_steps)) enter_production(ch);
                checkpoint_if_due(c, ch, true);  // forced final save (:1085)
            }
        }
        // Cross-chain statistics (:1117-1142).
        std::vector<Stats> cross(static_cast<size_t>(E + 1));
        for (auto& ch : chains)
            for (int o = c.order_min; o <= c.order_max; ++o) {
                const Stats& s = ch.result.order_signal[static_cast<size_t>(o)];
                if (s.count > 0) cross[static_cast<size_t>(o)].add(s.mean);
            }
        if (ratios) {
            const long double ref = cross[static_cast<size_t>(c.normalization_order)].mean;
            for (int o = 0; o <= E; ++o) ratios[o] = std::numeric_limits<double>::quiet_NaN();
            for (int o = c.order_min; o <= c.order_max; ++o)
                ratios[o] = ref == 0.0L ? std::numeric_limits<double>::quiet_NaN()
                                        : static_cast<double>(cross[static_cast<size_t>(o)].mean / ref);
        }
        if (results) {
            for (auto& ch : chains) {
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
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

}  // extern "C"

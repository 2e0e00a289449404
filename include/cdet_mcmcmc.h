/*
 * cdet_mcmcmc.h -- C-ABI of the set heat-bath MCMCMC driver (host, C++).
 *
 * The driver is the caller of the integrand path: it reproduces the
 * reference's sampler (include/cdet/mcmcmc_sampler.h) -- mask plan, fluxes,
 * weighted subset choice, seed regeneration, lambda adaptation, per-chain
 * host RNG (std::mt19937_64 seeded by splitmix64), and the text checkpoint
 * format CDET_MCMCMC_CHAIN_CHECKPOINT_V1 -- but advances all chains in
 * lockstep, so that one call evaluates the envelopes of every chain as one
 * batch.  Each chain's decisions depend only on its own RNG stream and its
 * own integrand values, so lockstep and the reference's one-thread-per-chain
 * schedule produce the same chains.
 */
#ifndef CDET_MCMCMC_H
#define CDET_MCMCMC_H

#include <stdint.h>

#include "cdet_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CDET_MCMCMC_MAX_ENVELOPE 48

/* Seed modes (include/cdet/mcmcmc_seed.h). */
#define CDET_SEED_UNIFORM 0
#define CDET_SEED_HISTOGRAM 1

/* Resume modes (MCMCMC_RESUME = off | auto | require). */
#define CDET_RESUME_OFF 0
#define CDET_RESUME_AUTO 1
#define CDET_RESUME_REQUIRE 2

/* MCMCMCConfig (include/cdet/mcmcmc_sampler.h:79-101). */
typedef struct cdet_mcmcmc_config {
    int32_t order_min;
    int32_t order_max;
    int32_t envelope_order;
    int32_t normalization_order;
    double lambdas[CDET_MCMCMC_MAX_ENVELOPE + 1]; /* indexed by order */
    int32_t lambda_warmup_steps;
    int32_t lambda_adapt_interval;
    double lambda_adapt_rate;
    int32_t production_steps;
    int32_t chain_count;
    uint64_t base_seed;
    int32_t seed_mode;
    int32_t checkpoint_enabled;
    uint64_t checkpoint_interval_steps;
    int32_t resume_mode;
    char checkpoint_dir[1024];
    /* Stop every chain after this many completed steps (warmup +
     * production) in this process, as an interrupted run would; 0 = run to
     * the end.  Used to rehearse kill-and-resume. */
    uint64_t stop_after_steps;
    /* Histogram seed (MCMCMCConfig histogram_* fields, mcmcmc_sampler.h:92-96;
     * lattice extents N1, N2 and BETA from the integrand parameters). */
    int32_t N1;
    int32_t N2;
    double beta;                     /* recorded in the histogram export only */
    int32_t histogram_learning_steps;
    int32_t histogram_tau_bins;
    double histogram_uniform_mix;
    double histogram_pseudo_count;
    int32_t histogram_learning_only; /* MCMCMC_SEED_HISTOGRAM_LEARNING_ONLY */
    char histogram_output_prefix[1024]; /* MCMCMC_SEED_HISTOGRAM_OUTPUT_PREFIX ("" = no export) */
    /* Host threads for the per-chain step work after each batched
     * evaluation (0 = min(16, hardware threads)); results do not depend on it. */
    int32_t host_threads;
} cdet_mcmcmc_config;

/* Fill the reference's MCMCMCConfig defaults (mcmcmc_sampler.h:79-101) plus
 * N1 = N2 = 1, beta = 1. */
void cdet_mcmcmc_config_defaults(cdet_mcmcmc_config* config);

/*
 * Factorized histogram seed (FactorizedHistogramSeed,
 * include/cdet/mcmcmc_seed.h:63-229): three row-stochastic tables indexed
 * [parent bin][child bin], row-major, over imaginary-time bins, lattice sites
 * (site = x * N2 + y) and the two replicas.  Arrays are owned by the caller.
 */
typedef struct cdet_histogram_seed {
    int32_t tau_bins;
    int32_t N1;
    int32_t N2;
    double uniform_mix;     /* learning settings, recorded by the export */
    double pseudo_count;
    double* tau_prob;       /* tau_bins * tau_bins */
    double* space_prob;     /* (N1 * N2) * (N1 * N2) */
    double replica_prob[4]; /* 2 * 2 */
} cdet_histogram_seed;

/*
 * Integrand callback: for `count` envelopes packed in h_vertices (envelope c
 * holds h_orders[c] vertices), write the all-subset integrands packed in
 * h_out (2^order doubles per envelope), as
 * evaluate_cdet_density_all_subset_integrands_double_into.  Returns 0 on
 * success.  cdet_eval_batch_host (below) and the CPU oracle provide it.
 */
typedef int (*cdet_batch_eval_fn)(void* user, const cdet_vertex* h_vertices, const int32_t* h_orders,
                                  int64_t count, double* h_out);

/*
 * The same evaluation in two steps over two slots (0, 1): submit takes a copy
 * of the batch, starts it and returns; wait blocks until the slot's batch is
 * done and writes its values (the layout of cdet_batch_eval_fn).  A slot holds
 * one batch at a time.  With both set, the driver keeps two groups of chains
 * in flight, so one group's heat-bath steps overlap the other's evaluation.
 * No counterpart in the reference, whose chains call the evaluator one at a
 * time from their own threads (include/cdet/mcmcmc_sampler.h:1107).
 */
typedef int (*cdet_batch_submit_fn)(void* user, int slot, const cdet_vertex* h_vertices, const int32_t* h_orders,
                                    int64_t count);
typedef int (*cdet_batch_wait_fn)(void* user, int slot, double* h_out);

/*
 * Optional allocator of the host value buffers handed to the evaluation
 * callbacks (user = the eval user).  A GPU backend gives pinned memory, which
 * its copies reach by DMA with no staging copy.  No counterpart in the
 * reference, whose evaluator fills a std::vector.
 */
typedef void* (*cdet_host_alloc_fn)(void* user, size_t bytes);
typedef void (*cdet_host_free_fn)(void* user, void* p);

/* Per-chain results (MCMCMCChainResult, include/cdet/mcmcmc_sampler.h:120-134). */
typedef struct cdet_mcmcmc_chain_result {
    uint64_t seed;
    uint64_t warmup_proposed_steps;
    uint64_t production_proposed_steps;
    uint64_t warmup_chosen_counts[CDET_MCMCMC_MAX_ENVELOPE + 1];
    uint64_t production_chosen_counts[CDET_MCMCMC_MAX_ENVELOPE + 1];
    double final_lambdas[CDET_MCMCMC_MAX_ENVELOPE + 1];
    double order_signal_mean[CDET_MCMCMC_MAX_ENVELOPE + 1];
    double order_signal_stderr[CDET_MCMCMC_MAX_ENVELOPE + 1];
    uint64_t order_signal_count[CDET_MCMCMC_MAX_ENVELOPE + 1];
    double retained_fraction_mean;
} cdet_mcmcmc_chain_result;

/*
 * The MCMCMC flow of the reference's main (src/main.cpp:999-1107): validate
 * the configuration; with seed_mode = CDET_SEED_HISTOGRAM learn the histogram
 * seed (cdet_histogram_learn), export it when histogram_output_prefix is set,
 * and stop there when histogram_learning_only is set (results untouched);
 * then run all chains (run_mcmcmc_chains_parallel,
 * include/cdet/mcmcmc_sampler.h:1090-1145) with that seed.
 * results: chain_count entries.  ratios: envelope_order + 1 entries,
 * cross-chain order-signal mean ratio to the normalization order.
 * trace (may be NULL): chain_count * trace_stride chosen masks, step k of chain
 * c (warmup then production, steps of this process only) at
 * trace[c * trace_stride + k]; steps beyond trace_stride are not recorded.
 */
int cdet_mcmcmc_run(const cdet_mcmcmc_config* config, cdet_batch_eval_fn eval, void* eval_user,
                    cdet_mcmcmc_chain_result* results, double* ratios, uint64_t* trace, int64_t trace_stride);

/*
 * Run all chains with a given seed kernel (run_mcmcmc_chains_parallel with
 * seed_kernel, mcmcmc_sampler.h:1090): seed = NULL is the uniform seed, else
 * the histogram seed.  config->seed_mode is still validated and written to
 * checkpoints as is, so seed = NULL with seed_mode = histogram is the
 * reference's MCMCMC_COMPARE_UNIFORM_SEED rerun (src/main.cpp:1109-1120).
 */
int cdet_mcmcmc_run_seeded(const cdet_mcmcmc_config* config, const cdet_histogram_seed* seed,
                           cdet_batch_eval_fn eval, void* eval_user, cdet_mcmcmc_chain_result* results,
                           double* ratios, uint64_t* trace, int64_t trace_stride);

/*
 * Chains with their own envelope orders (no counterpart in the reference,
 * whose chains share MCMCMC_ENVELOPE_ORDER): chain k runs the reference's
 * sampler with envelope_order = envelope_orders[k] and order_max =
 * min(config->order_max, envelope_orders[k] - 1); its checkpoint records those
 * two values.  Each chain's settings must pass the reference's validation
 * (so normalization_order < every envelope order).  ratios (max envelope
 * order + 1 entries): for each order, the mean over the chains that measure it
 * of the chain's order-signal mean divided by its normalization-order signal
 * mean (chains of different envelopes normalise by different partition sums,
 * so their raw signals are not averaged).  seed as in cdet_mcmcmc_run_seeded.
 */
int cdet_mcmcmc_run_envelopes(const cdet_mcmcmc_config* config, const cdet_histogram_seed* seed,
                              const int32_t* envelope_orders, cdet_batch_eval_fn eval, void* eval_user,
                              cdet_mcmcmc_chain_result* results, double* ratios, uint64_t* trace,
                              int64_t trace_stride);

/*
 * Learn the histogram seed (learn_factorized_histogram_seed,
 * include/cdet/mcmcmc_sampler.h:607-671) from histogram_learning_steps
 * uniform envelopes drawn from one std::mt19937_64 stream, evaluated in
 * batches through `eval`.  seed->tau_prob and seed->space_prob must hold
 * tau_bins^2 and (N1*N2)^2 doubles; the scalar fields are set from config.
 * total_pair_weight (may be NULL) receives the accumulated directed-pair weight.
 */
int cdet_histogram_learn(const cdet_mcmcmc_config* config, cdet_batch_eval_fn eval, void* eval_user,
                         cdet_histogram_seed* seed, double* total_pair_weight);

/* Write <prefix>_tau.csv, _space.csv, _replica.csv and _metadata.txt
 * (export_factorized_histogram_seed, src/main.cpp:704-767). */
int cdet_histogram_export(const cdet_histogram_seed* seed, double beta, const char* prefix);

/* Conditional generation probabilities of all 2^E subsets of an envelope
 * (compute_conditional_generation_probabilities, mcmcmc_sampler.h:252-330),
 * double recursion; seed = NULL is the uniform seed (all ones). */
int cdet_seed_conditional_probabilities(const cdet_histogram_seed* seed, const cdet_vertex* vertices,
                                        int envelope_order, double* out);

/* Draw out[i] = sample_child_from_parent(parents[i]) for i = 0..n-1 from one
 * std::mt19937_64(rng_seed) stream (mcmcmc_seed.h:51-55, :106-128). */
int cdet_seed_sample_children(const cdet_histogram_seed* seed, const cdet_vertex* parents, int64_t n,
                              uint64_t rng_seed, cdet_vertex* out);

/* Density q(child | parent) of the seed kernel (mcmcmc_seed.h:57-60, :130-150). */
int cdet_seed_density(const cdet_histogram_seed* seed, const cdet_vertex* child, const cdet_vertex* parent,
                      double* out);

/* Message of the last failing driver call on this host thread. */
const char* cdet_mcmcmc_last_error(void);

/*
 * Conditional-generation callback: for `count` envelopes packed as for
 * cdet_batch_eval_fn, write the histogram seed's conditional generation
 * probabilities of every subset, packed 2^order per envelope, bit-identical
 * to cdet_seed_conditional_probabilities.  Returns 0 on success.
 * cdet_seed_cond_batch_host (below) provides it on the GPU.
 */
typedef int (*cdet_batch_cond_fn)(void* user, const cdet_histogram_seed* seed, const cdet_vertex* h_vertices,
                                  const int32_t* h_orders, int64_t count, double* h_out);

/* Where a run evaluates: the integrand batch callback, and optionally the
 * conditional-generation callback of the histogram seed (NULL: the driver's
 * own host recursion, on host threads).  Either choice gives the same chains. */
typedef struct cdet_mcmcmc_backend {
    cdet_batch_eval_fn eval;
    void* eval_user;
    cdet_batch_cond_fn cond;
    void* cond_user;
    cdet_batch_submit_fn submit;  /* optional, with wait: two groups of chains in flight (user = eval_user) */
    cdet_batch_wait_fn wait;
    cdet_host_alloc_fn host_alloc; /* optional, with host_free: allocator of the value buffers (user = eval_user) */
    cdet_host_free_fn host_free;
} cdet_mcmcmc_backend;

/* cdet_mcmcmc_run_envelopes with a backend; envelope_orders may be NULL (every
 * chain at config->envelope_order, the reference's setting). */
int cdet_mcmcmc_run_backend(const cdet_mcmcmc_config* config, const cdet_histogram_seed* seed,
                            const int32_t* envelope_orders, const cdet_mcmcmc_backend* backend,
                            cdet_mcmcmc_chain_result* results, double* ratios, uint64_t* trace,
                            int64_t trace_stride);

/*
 * GPU implementation of cdet_batch_eval_fn over a context (user = cdet_ctx*):
 * host buffers in and out, one batched plan per distinct order list (cached),
 * mode from cdet_ctx_set_batch_mode (default CDET_MODE_EXACT, bit-identical
 * to the plain form).  Synchronous.
 */
int cdet_eval_batch_host(void* ctx, const cdet_vertex* h_vertices, const int32_t* h_orders, int64_t count,
                         double* h_out);
/* GPU implementation of cdet_batch_submit_fn / cdet_batch_wait_fn over a
 * context: each slot has its own plan cache, stream and pinned staging, so the
 * two slots' batches run concurrently; same values as cdet_eval_batch_host. */
int cdet_eval_batch_submit(void* ctx, int slot, const cdet_vertex* h_vertices, const int32_t* h_orders,
                           int64_t count);
int cdet_eval_batch_wait(void* ctx, int slot, double* h_out);
/* cdet_host_alloc_fn / cdet_host_free_fn over a context: pinned host memory
 * that the context remembers.  An h_out of cdet_eval_batch_host inside such
 * a buffer receives the values by DMA, without the copy out of the context's
 * staging.  Freed by cdet_host_free or with the context. */
void* cdet_host_alloc(void* ctx, size_t bytes);
void cdet_host_free(void* ctx, void* p);
int cdet_ctx_set_batch_mode(cdet_ctx* ctx, int mode);
/* Green-pair reuse of those plans (cdet_batch_plan_reuse_pairs), on by
 * default: consecutive calls are steps of the same chains, and a chain's
 * retained vertices keep their Green pairs, as the reference's pair cache
 * does (include/cdet/integrand_replica_randomly_sampled.h:664-696).  Values
 * are the same either way. */
int cdet_ctx_set_pair_reuse(cdet_ctx* ctx, int enable);

/* GPU implementation of cdet_batch_cond_fn over a context (user = cdet_ctx*):
 * one workgroup per envelope walks the popcount layers of the subset
 * recursion with the probabilities in LDS (envelopes up to CDET_MAX_ORDER).
 * The seed tables are uploaded when they change.  Synchronous. */
int cdet_seed_cond_batch_host(void* ctx, const cdet_histogram_seed* seed, const cdet_vertex* h_vertices,
                              const int32_t* h_orders, int64_t count, double* h_out);

#ifdef __cplusplus
}
#endif

#endif /* CDET_MCMCMC_H */

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
} cdet_mcmcmc_config;

/*
 * Integrand callback: for `count` envelopes packed in h_vertices (envelope c
 * holds h_orders[c] vertices), write the all-subset integrands packed in
 * h_out (2^order doubles per envelope), as
 * evaluate_cdet_density_all_subset_integrands_double_into.  Returns 0 on
 * success.  cdet_eval_batch_host (below) and the CPU oracle provide it.
 */
typedef int (*cdet_batch_eval_fn)(void* user, const cdet_vertex* h_vertices, const int32_t* h_orders,
                                  int64_t count, double* h_out);

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
 * Run all chains (run_mcmcmc_chains_parallel, include/cdet/mcmcmc_sampler.h:1090-1145).
 * results: chain_count entries.  ratios: envelope_order + 1 entries,
 * cross-chain order-signal mean ratio to the normalization order.
 * trace (may be NULL): chain_count * trace_stride chosen masks, step k of chain
 * c (warmup then production, steps of this process only) at
 * trace[c * trace_stride + k]; steps beyond trace_stride are not recorded.
 */
int cdet_mcmcmc_run(const cdet_mcmcmc_config* config, cdet_batch_eval_fn eval, void* eval_user,
                    cdet_mcmcmc_chain_result* results, double* ratios, uint64_t* trace, int64_t trace_stride);

/* Message of the last failing driver call on this host thread. */
const char* cdet_mcmcmc_last_error(void);

/*
 * GPU implementation of cdet_batch_eval_fn over a context (user = cdet_ctx*):
 * host buffers in and out, one batched plan per distinct order list (cached),
 * mode from cdet_ctx_set_batch_mode (default CDET_MODE_EXACT, bit-identical
 * to the plain form).  Synchronous.
 */
int cdet_eval_batch_host(void* ctx, const cdet_vertex* h_vertices, const int32_t* h_orders, int64_t count,
                         double* h_out);
int cdet_ctx_set_batch_mode(cdet_ctx* ctx, int mode);

#ifdef __cplusplus
}
#endif

#endif /* CDET_MCMCMC_H */

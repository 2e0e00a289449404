/*
 * cdet_hip.h -- C-ABI of the MI355X (gfx950) connected-determinant (CDet)
 * integrand path.
 *
 * The reference exposes this path only as in-process C++ templates driven by
 * process globals.  Every entry point below replaces one reference call and
 * takes its implicit inputs (globals in include/cdet/parameters.h:11-88 of the
 * reference) explicitly through cdet_params / cdet_ctx.  All entry points are
 * plain C: pointers, sizes and status codes, no C++ or torch types.
 *
 * Status codes: 0 = success, negative = error; the message of the most recent
 * error on the calling host thread is returned by cdet_last_error().  The
 * messages of errors that the reference raises as std::runtime_error carry the
 * reference's text so that a host shim can rethrow them unchanged.
 *
 * Threading.  The reference calls the integrand concurrently from one
 * std::async thread per chain (include/cdet/mcmcmc_sampler.h:1104-1110), each
 * with thread_local workspaces.  Here:
 *  - cdet_eval_single / cdet_eval_all_subsets may be called from any number of
 *    host threads at once on one context: each call takes a workspace (stream
 *    and device buffers) of the context's pool, or makes one, and returns it
 *    when it has finished; results do not depend on the interleaving.
 *  - cdet_green_pair_eval calls from several threads are serialised on the
 *    context's lookup scratch (the later call's launches wait on the device
 *    for the earlier call's).
 *  - cdet_eval_batch_host holds a context lock for the whole call;
 *    cdet_eval_batch_submit / _wait on one slot belong to one thread.
 *  - A cdet_batch_plan is used by one thread at a time; different plans of a
 *    context may run concurrently on different streams.
 *  - Tables (upload, alias, build, chunk loads) change only while no
 *    evaluation runs, as the reference fills its globals before the chains
 *    start (include/cdet/load_data.h:119-178).
 *  - Counters are atomic; call counts are exact, the fallback count covers
 *    work that has completed on the device.
 */
#ifndef CDET_HIP_H
#define CDET_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define CDET_OK 0
#define CDET_ERR_ARGUMENT (-1)      /* bad argument / shape */
#define CDET_ERR_ORDER (-2)         /* order <= 1 where the reference throws */
#define CDET_ERR_NONFINITE (-3)     /* "Value is not finite" throw sites */
#define CDET_ERR_TABLES (-4)        /* table missing / mismatched */
#define CDET_ERR_DEVICE (-5)        /* HIP runtime error */
#define CDET_ERR_CAPACITY (-6)      /* order or size beyond the compiled maximum */
#define CDET_ERR_SOLVE (-7)         /* producer: singular / not-real entry */
#define CDET_ERR_IO (-8)            /* chunk directory: missing file, bad size or manifest */

/* Largest envelope order accepted by the integrand entry points. */
#define CDET_MAX_ORDER 14

/*
 * One interaction vertex as four unit coordinates in [0,1): imaginary time,
 * x, y, replica.  Layout-identical to the reference's UnitVertex
 * (include/cdet/cdet_integrand_evaluator.h:16-21).
 */
typedef struct cdet_vertex {
    double tau;
    double x;
    double y;
    double replica;
} cdet_vertex;

/*
 * Scalar inputs of the integrand.  Replaces the process globals read by the
 * reference hot path (include/cdet/parameters.h:11-24) and the runtime scalar
 * cache (include/cdet/integrand_replica_randomly_sampled.h:98-119).
 *
 * The three derived scalars (scale, diag_up_shift, diag_down_shift) are the
 * double roundings of sqrt(U*N1*N2*beta), alpha_down/U and alpha_up/U.  The
 * reference forms them in double-double from the decimal parameter text;
 * cdet_params_derive() forms them from the double fields with 113-bit
 * arithmetic, and a caller that holds the decimal text may overwrite them.
 */
typedef struct cdet_params {
    int32_t N1;
    int32_t N2;
    int32_t cheb_nu;              /* coefficient count along u (degree+1) */
    int32_t cheb_nv;              /* coefficient count along v (degree+1) */
    double beta;
    double U;
    double mu;
    double alpha_up;
    double alpha_down;
    double branch_epsilon;
    double tol_fpm;
    double fpm_global_shift;      /* < 0: automatic shift max|A|/n^2 */
    double ztransform_epsilon0;   /* <= 0: automatic (2e-16 for double) */
    double ztransform_safety;     /* 0 disables the round-off filter */
    double scale;                 /* derived: sqrt(U*N1*N2*beta) */
    double diag_up_shift;         /* derived: alpha_down/U */
    double diag_down_shift;       /* derived: alpha_up/U */
} cdet_params;

/* Table kinds (include/cdet/chunked_load_data.h:29-34 of the reference). */
#define CDET_TABLE_TWISTED_UP 0
#define CDET_TABLE_TWISTED_DOWN 1
#define CDET_TABLE_ORDINARY_UP 2
#define CDET_TABLE_ORDINARY_DOWN 3

/* Evaluation modes of cdet_eval_batch. */
#define CDET_MODE_EXACT 0     /* FP64 Clenshaw in the reference's order: bit-identical */
#define CDET_MODE_MFMA64 1    /* FP64 MFMA coefficient x basis contraction, tree-order recursion sums */
#define CDET_MODE_SPLIT32 2   /* split-FP32 MFMA (hi/lo, three products), FP64 combine, tree-order sums */

typedef struct cdet_counters {
    uint64_t integrand_calls;     /* include/cdet/runtime_counters.h:14 */
    uint64_t fpm_fallbacks;       /* include/cdet/runtime_counters.h:15 */
    uint64_t green_pair_cache_hits;
    uint64_t green_pair_cache_misses;
} cdet_counters;

typedef struct cdet_ctx cdet_ctx;

/* Message of the last failing call on this host thread ("" if none). */
const char* cdet_last_error(void);

/* Fill scale / diag_up_shift / diag_down_shift from the double fields. */
int cdet_params_derive(cdet_params* params);

/*
 * Create a context on the current HIP device.  The context owns device copies
 * of the tables and all scratch; nothing is allocated on the per-call path
 * once the scratch has grown to the largest order seen.
 */
int cdet_ctx_create(const cdet_params* params, cdet_ctx** out_ctx);
int cdet_ctx_destroy(cdet_ctx* ctx);

/*
 * Upload one resident table kind from host memory.  Layout as the reference's
 * PiecewiseChebyshevTable (include/cdet/replica_chebyshev_tables.h:25-33):
 * coeffs [sitepair][a][b] with b fastest, sitepair count 4*(N1*N2)^2
 * (include/cdet/replica_chebyshev_tables.h:360-400); corners [sitepair].
 * Replaces the resident globals g_table_* (include/cdet/parameters.h:81-84).
 */
int cdet_tables_upload(cdet_ctx* ctx, int kind,
                       const double* upper_coeffs, const double* lower_coeffs,
                       const double* upper_corners, const double* lower_corners);

/* Make table kind dst share the device storage of kind src
 * (include/cdet/load_data.h:70-91: one table reused when ALPHA_UP == ALPHA_DOWN). */
int cdet_tables_alias(cdet_ctx* ctx, int dst_kind, int src_kind);

/*
 * Lattice description for the GPU table producer.  mask_kind follows
 * include/cdet/replica_free_greens.h:20-26 (0 empty, 1 checkerboard,
 * 2 first_row, 3 first_column, 4 origin_square) plus 5 = left_columns
 * (sites with x < mask_param), used for half-strip and half-lattice regions.
 */
typedef struct cdet_lattice {
    int32_t pbc;
    int32_t mask_kind;
    int32_t mask_param;           /* origin_square radius / left_columns width */
    int32_t node_chunk;           /* Chebyshev nodes per producer launch; <= 0: all */
    int32_t ordinary_layout;      /* CDET_ORDINARY_RESIDENT or CDET_ORDINARY_CHUNKED */
} cdet_lattice;

/* Ordinary-table layouts.  RESIDENT: the replica (0,0) block in site-major
 * order, as the reference's resident load_data builds it
 * (include/cdet/load_data.h:39-43, replica_chebyshev_tables.h:697-793).
 * CHUNKED: every replica block in the consistent sitepair layout, as its
 * chunked workflow builds it (include/cdet/chunked_load_data.h:594-604).
 * Lookups read the same index in both; the two agree by the lattice's x<->y
 * reflection symmetry only. */
#define CDET_ORDINARY_RESIDENT 0
#define CDET_ORDINARY_CHUNKED 1

/*
 * Build the twisted and ordinary tables on the GPU in FP64 (the reference's
 * load_data(), include/cdet/load_data.h:49-99 with build_green_function_table,
 * include/cdet/replica_chebyshev_tables.h:592-793).  When alpha_up ==
 * alpha_down one table per class is built and aliased.
 */
int cdet_tables_build(cdet_ctx* ctx, const cdet_lattice* lattice);

/* Copy a device table kind back to host buffers (same layout as upload). */
int cdet_tables_download(cdet_ctx* ctx, int kind,
                         double* upper_coeffs, double* lower_coeffs,
                         double* upper_corners, double* lower_corners);

/*
 * Copy selected coefficient blocks of a device table kind back to host
 * buffers: block i of each branch (nu*nv doubles) is the table's block
 * sitepairs[i], written at upper/lower + i*nu*nv; corner i at ucorner/lcorner
 * + i.  Any output may be NULL.  The partial form of cdet_tables_download for
 * tables too large to copy whole (the 10x10, 192-coefficient tables hold
 * 23.6 GB per kind); no counterpart in the reference.
 */
int cdet_tables_download_blocks(cdet_ctx* ctx, int kind, const int64_t* sitepairs, int64_t nblocks,
                                double* upper, double* lower, double* ucorner, double* lcorner);

/*
 * Single-order integrand C(n) of all n vertices, twisted - 2*ordinary combined
 * in double-double and rounded.  Replaces
 * evaluate_cdet_density_integrand_double
 * (include/cdet/cdet_integrand_evaluator.h:542-560).  n >= 1.
 */
int cdet_eval_single(cdet_ctx* ctx, const cdet_vertex* vertices, int n, double* out_value);

/*
 * All-subset integrand: out[mask] for every mask of the E-vertex envelope
 * (out has 2^E doubles, out[0] = 0).  Replaces
 * evaluate_cdet_density_all_subset_integrands_double_into
 * (include/cdet/cdet_integrand_evaluator.h:562-583).  E >= 2.
 * Plain form: one configuration per launch sequence, FP64, bit-identical to
 * the reference arithmetic.
 */
int cdet_eval_all_subsets(cdet_ctx* ctx, const cdet_vertex* vertices, int envelope_order,
                          double* out_values);

/*
 * All-subset integrand of one configuration that allocates nothing: the same
 * kernels and values as cdet_eval_all_subsets (bit for bit), on DEVICE buffers
 * (d_vertices: cdet_vertex[E]; d_out: double[2^E]; d_status: int32[1] or
 * NULL, bit 0 = a non-finite value, bit 1 = FPM fallback taken),
 * asynchronous on `stream` (hipStream_t; NULL = the context stream), with no
 * host synchronisation.  All other device memory it needs (Green values,
 * minors, queues: at most 3.1 MB, for E = 14) lies in one 4 MiB file-scope
 * device array of the library, which each GPU the library runs on holds until
 * the process ends.  E from 2 to CDET_MAX_ORDER; larger requests are rejected
 * with CDET_ERR_CAPACITY.  Every call on a GPU shares that GPU's array: calls
 * that overlap in time on one GPU (different streams, or several contexts)
 * race and must be serialised by the caller.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).
 */
int cdet_eval_all_subsets_noalloc(cdet_ctx* ctx, const cdet_vertex* d_vertices, int envelope_order,
                                  double* d_out, int32_t* d_status, void* stream);

/*
 * Batched evaluation.  A plan fixes the order of every configuration
 * (h_orders: host int32[count]); configurations are packed in the given order:
 * configuration c owns vertices [sum_{c'<c} n_c', ... + n_c) and, for the
 * all-subset kind, outputs [sum_{c'<c} 2^n_c', ... + 2^n_c) (out[mask] as
 * cdet_eval_all_subsets); for the single-order kind one output per
 * configuration.  mode is one of CDET_MODE_*; eval_kind 0 = all subsets
 * (orders 2..CDET_MAX_ORDER), 1 = single order (orders 1..CDET_MAX_ORDER),
 * 2 = conditioning scale of every all-subset value (same layout as kind 0):
 * |prefactor| * (|a[mask]| + sum of |subtracted terms|) of the subset
 * recursion, summed as twisted + 2*ordinary; a value computed with relative
 * perturbations of size e in its inputs can differ from the exact one by
 * about e times this scale, however small the value itself; 3 = first-order
 * error bound of every all-subset value (same layout as kind 0): running
 * error analysis of the reference's algorithm (Schur-complement minors with
 * shift removal, pivoted fallback, subset recursion with its round-off
 * filter, assembly) under an absolute error of each table value of at most
 * (nu + nv) u sum|c_ab| over its coefficient block plus one rounding of
 * u = 2^-53 per operation.  Two evaluations of a configuration whose table
 * values satisfy that (any FP64 summation order of the Chebyshev sums) differ
 * by at most about twice this bound.
 * Batched form of the reference's per-chain calls
 * (include/cdet/mcmcmc_sampler.h:538, :624-625; cdet_integrand_evaluator.h:542-583).
 */
typedef struct cdet_batch_plan cdet_batch_plan;
int cdet_batch_plan_create(cdet_ctx* ctx, const int32_t* h_orders, int64_t count, int mode, int eval_kind,
                           cdet_batch_plan** out_plan);
int cdet_batch_plan_destroy(cdet_batch_plan* plan);  /* before its context's cdet_ctx_destroy */
int64_t cdet_batch_out_size(const cdet_batch_plan* plan);     /* doubles written by one run */
int64_t cdet_batch_vertex_count(const cdet_batch_plan* plan); /* vertices read by one run */
int64_t cdet_batch_query_count(const cdet_batch_plan* plan);  /* Green-pair slots per run (2 sum n^2) */
/* Lookups the plan's last run sent to the table contraction and the tiles it
 * cut them into (waits for that run).  Corner pairs, reused pairs and, in the
 * FP64 matrix-core mode, the diagonal pairs (evaluated from row sums) are not
 * among them.  No counterpart in the reference (measurement). */
int cdet_batch_contracted(cdet_batch_plan* plan, int64_t* lookups, int64_t* tiles);
/* Streams a run spreads the plan's order buckets over (1 = the run's stream
 * only; up to 4 for batches of several orders, forked from and joined back to
 * the run's stream).  No counterpart in the reference. */
int cdet_batch_plan_streams(const cdet_batch_plan* plan);

/*
 * Green-pair reuse across the plan's runs (enable != 0; off by default):
 * each configuration's vertices are compared, bit for bit, with those of the
 * plan's previous run, and a Green pair of two vertices that were present
 * (at any position) copies its previous value instead of being looked up
 * again, as the reference's GreenPairMatrixCache does for consecutive calls
 * on one thread (include/cdet/integrand_replica_randomly_sampled.h:57-86,
 * :664-696).  The values are those a lookup computes, in every mode.  Values
 * of an earlier table generation are never reused.  Ignored for the
 * error-bound kind.  Counted in cdet_counters.green_pair_cache_hits/misses.
 */
int cdet_batch_plan_reuse_pairs(cdet_batch_plan* plan, int enable);

/*
 * Run a plan on DEVICE buffers: d_vertices (cdet_vertex[vertex_count]),
 * d_out (double[out_size]), d_status (int32[count] or NULL: bit 0 = a
 * non-finite value, bit 1 = FPM fallback taken).  Asynchronous on `stream`
 * (hipStream_t; NULL = the context stream).  With use_graph != 0 the launch
 * sequence is captured once per (buffers, stream) into a hipGraph and
 * replayed.  Status reports launch errors only.
 */
int cdet_batch_run(cdet_batch_plan* plan, const cdet_vertex* d_vertices, double* d_out, int32_t* d_status,
                   int use_graph, void* stream);

/*
 * Eager run with HIP events between the three stages, on `stream`: stage_ms[0]
 * = lookup planning and grouping, [1] = coefficient x basis contraction,
 * [2] = minors, subset recursion and assembly.  The contraction (which reads
 * the grouped lookups and writes the pair values, the same values each time)
 * is launched CDET_PROFILE_CONTRACT_REPS times back to back between its two
 * events and [1] is the mean, so that the events' own overhead does not enter
 * the kernel's duration.  Synchronises.  For measurement; results are
 * identical to cdet_batch_run.
 */
#define CDET_PROFILE_CONTRACT_REPS 10
int cdet_batch_profile(cdet_batch_plan* plan, const cdet_vertex* d_vertices, double* d_out, int32_t* d_status,
                       void* stream, float* stage_ms);

/* One-shot batched all-subset evaluation: plan, run, wait, free. */
int cdet_eval_batch(cdet_ctx* ctx, const cdet_vertex* d_vertices, const int32_t* h_orders, int64_t count,
                    double* d_out, int32_t* d_status, int mode, void* stream);

/*
 * Table evaluation only: (G_up, G_down) for `count` lookups, device pointers.
 * Each query is (kind, replica_alpha, replica_beta, x1, y1, x2, y2) packed in
 * int32[7] and (tau1, tau2) in double[2]; upper branch iff tau1 > tau2.
 * Replaces prepare_green_lookup_descriptor + evaluate_green_pair_descriptor
 * (include/cdet/g_funcs.h:609-653, :793-818).  out holds 2 doubles per query.
 */
int cdet_green_pair_eval(cdet_ctx* ctx, const int32_t* query_ints, const double* query_taus,
                         int64_t count, double* out, int mode, void* stream);

/*
 * cdet_green_pair_eval with HIP events between its stages, on `stream`:
 * stage_ms[0] = lookup planning and grouping, [1] = the coefficient x basis
 * contraction (mean of CDET_PROFILE_CONTRACT_REPS launches back to back),
 * [2] = the interleave into out.  Synchronises.  No counterpart in the
 * reference; for measurement, with the same results as cdet_green_pair_eval.
 */
int cdet_green_pair_profile(cdet_ctx* ctx, const int32_t* query_ints, const double* query_taus,
                            int64_t count, double* out, int mode, void* stream, float* stage_ms);

/*
 * Launch-shape tuning of a context.  No counterpart in the reference, whose
 * only work-split knob is its integrand thread pool
 * (include/cdet/integrand_parallel.h:34-154).  The package's autotuner
 * (tune.py) times these choices per workload on the device and writes a
 * table that callers apply with cdet_ctx_set_tuning; plans read the values in
 * effect when they are created.  Every choice gives bit-identical values.
 * A field <= 0 keeps the built-in default (in parentheses).
 */
typedef struct cdet_tuning {
    int32_t contract_wg_per_cu;   /* persistent contraction workgroups per CU, 1..8 (8) */
    int32_t contract_tile;        /* lookups per FP64 matrix-core tile, 16 or 64 (by batch size) */
    int32_t pair_min_per_cu;      /* both classes in one workgroup when every order bucket holds at
                                     least this many configurations per CU (2); a large value turns
                                     pairing off */
    int32_t max_lanes;            /* streams a plan's order buckets may spread over, 1..4 (4) */
    int32_t minors_threads[CDET_MAX_ORDER + 1];    /* per order: threads per workgroup of the general
                                                      minors kernels, 256 or 1024 (1024 from order 13) */
    int32_t group16_min_terms[CDET_MAX_ORDER + 1]; /* per order: a popcount layer of at most 16 masks
                                                      whose chains have at least this many terms runs
                                                      on 16 lanes per mask (64) */
    int32_t group64_min_terms[CDET_MAX_ORDER + 1]; /* per order: a single chain of at least this many
                                                      terms runs on a whole wavefront (128) */
    int32_t fused_minors_max_order; /* paired order buckets up to this order (at most 10) run principal
                                       minors and subset recursion in one workgroup per configuration,
                                       the minors kept in LDS; 1 = none (1) */
} cdet_tuning;
int cdet_ctx_set_tuning(cdet_ctx* ctx, const cdet_tuning* tuning);
int cdet_ctx_get_tuning(cdet_ctx* ctx, cdet_tuning* out); /* the values in effect */

/*
 * Order of the subset recursion's sums in the non-exact modes.  No counterpart
 * in the reference, whose sums run serially
 * (include/cdet/integrand_replica_randomly_sampled.h:326-343).
 * CDET_SUM_TREE (the default): plans in CDET_MODE_MFMA64 and CDET_MODE_SPLIT32
 * sum the terms of each connected value on several lanes and add the lanes'
 * sums pairwise, the lanes per mask fixed per order and popcount layer, so the
 * values do not depend on the launch shapes of cdet_ctx_set_tuning.
 * CDET_SUM_REFERENCE: the reference's serial order in every mode.
 * CDET_MODE_EXACT and the plain entry points always sum in the reference's
 * order.  Plans read the setting when they run; do not change it while runs
 * on the context are in flight.
 */
#define CDET_SUM_TREE 0
#define CDET_SUM_REFERENCE 1
int cdet_ctx_set_sum_order(cdet_ctx* ctx, int order);

/* Runtime counters (include/cdet/runtime_counters.h:7-51). */
int cdet_get_counters(cdet_ctx* ctx, cdet_counters* out);
int cdet_reset_counters(cdet_ctx* ctx);

/* Context stream (hipStream_t) on which the library launches by default. */
void* cdet_ctx_stream(cdet_ctx* ctx);

/* A stream (hipStream_t) on a hardware queue of its own, over every CU of the
 * context's device (call with that device current).  HIP multiplexes ordinary
 * streams over a pool of GPU_MAX_HW_QUEUES hardware queues (4 by default), so
 * two batches run on two ordinary streams can still execute one after the
 * other; plans run concurrently overlap only on streams of their own queues.
 * The library's own streams (context stream, host slots, order lanes) are
 * ordinary non-blocking streams, which measured faster for what they carry.  Like a hipStreamCreate stream it synchronizes with the
 * null stream.  Destroy it with cdet_stream_destroy before the context.  No
 * counterpart in the reference, whose chains are host threads
 * (include/cdet/mcmcmc_sampler.h:1104-1110). */
int cdet_stream_create(cdet_ctx* ctx, void** out_stream);
int cdet_stream_destroy(cdet_ctx* ctx, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CDET_HIP_H */

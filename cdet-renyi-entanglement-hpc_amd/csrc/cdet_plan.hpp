// cdet_plan.hpp -- batched evaluation plans (included by cdet_kernels.hip).
//
// A plan fixes the order of every configuration of a batch.  It owns the
// per-configuration offsets, the order buckets and all scratch, so that a run
// is a fixed launch sequence with no allocation and no host synchronisation;
// the sequence is captured once into a hipGraph and replayed.

constexpr double kForkMinUs = 40.0;  // bucket work beyond the longest bucket worth a fork

struct cdet_batch_plan {
    cdet_ctx* ctx = nullptr;
    int contract_mode = CDET_MODE_EXACT;
    int eval_kind = 0;  // 0 all-subset, 1 single-order
    int64_t count = 0;
    int64_t nq = 0;         // 2 * sum n^2 query slots
    int64_t out_total = 0;  // doubles written
    int64_t vert_total = 0;
    int32_t nkeys = 0;
    int64_t max_tiles = 0;
    int tile = 64;
    std::vector<int> bucket_off, bucket_cnt;  // by order
    int32_t* d_order = nullptr;
    int64_t* d_vbase = nullptr;
    int64_t* d_pbase = nullptr;
    int64_t* d_obase = nullptr;
    int32_t* d_lists = nullptr;
    double* d_gup = nullptr;
    double* d_gdn = nullptr;
    int32_t* d_qkey = nullptr;
    int32_t* d_qrank = nullptr;
    double* d_qu = nullptr;
    double* d_qv = nullptr;
    int32_t* d_hist = nullptr;
    int32_t* d_keyoff = nullptr;
    int2* d_wgtot = nullptr;  // per-workgroup totals of the key scan
    int32_t* d_tkey = nullptr;
    int32_t* d_tstart = nullptr;
    int32_t* d_tlen = nullptr;
    int32_t* d_ntiles = nullptr;
    double* d_su = nullptr;
    double* d_sv = nullptr;
    int32_t* d_sdest = nullptr;
    double* d_out_ord = nullptr;
    double* d_single = nullptr;
    int32_t* d_fb_list = nullptr;
    int32_t* d_fb_count = nullptr;  // one queue length per order
    int32_t* d_defer_list = nullptr;   // configurations k_minors_pair defers (bucket offsets as d_lists)
    int32_t* d_defer_count = nullptr;  // two words per order
    double* d_eup = nullptr;        // error-bound kind: table-value error scale per Green pair
    double* d_edn = nullptr;
    // Green-pair reuse across runs (cdet_batch_plan_reuse_pairs)
    int reuse = 0;
    double* d_prev_verts = nullptr;  // vertices of the previous run, same packing
    int32_t* d_prev_ok = nullptr;    // per configuration: previous vertices and values valid
    uint64_t reuse_tables = 0;       // table generation the previous values belong to
    Scratch scratch;                // minors global scratch of lane 0, sized at creation
    int contract_grid = 0;
    // Minors runs, highest order first: one per order bucket, except that
    // the buckets of orders 2 .. 7 and the small ones of orders 8 .. 10 form a
    // single run of mixed orders when more than one of them is non-empty
    // (modes 0 - 2), which saves the launch chains of the small buckets.
    struct MinorRun {
        int n;    // the run's order (mixed: the highest)
        int off;  // first list entry
        int cnt;  // configurations
        bool mixed;
    };
    std::vector<MinorRun> runs;
    // Minors runs spread over streams: run r runs on lane run_lane[r]
    // (lane 0 = the run's stream, lane k > 0 = aux[k - 1], forked from and
    // joined back to the run's stream by events); each lane has its own
    // minors scratch, since its runs run one after another.  The aux
    // streams belong to the context (or to the submit slot the plan serves)
    // and are shared by its plans.
    // Both classes of a configuration in one k_cdet workgroup, combined in
    // place (no k_combine_all pass over the outputs): when every order bucket
    // holds at least pair_min_per_cu (tuning) configurations per CU, so that the longer
    // workgroups still fill the device.
    bool paired = false;
    int lanes = 1;
    int max_lanes = kMinorLanes;
    std::vector<int> run_lane;
    Scratch lane_scratch[kMinorLanes];  // [0] unused (lane 0 uses scratch)
    hipStream_t aux[kMinorLanes - 1] = {};
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join[kMinorLanes - 1] = {};
    // graph cache
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint64_t g_tables = 0;  // table generation the graph was captured with
    const void* g_verts = nullptr;
    void* g_out = nullptr;
    void* g_status = nullptr;
    hipStream_t g_stream = nullptr;
    // recorded on the run's stream after each run's last launch; the plan is
    // freed only once it has completed
    hipEvent_t done = nullptr;
    bool ran = false;
};

static int plan_mark_done(cdet_batch_plan* p, hipStream_t s) {
    HIP_TRY(hipEventRecord(p->done, s));
    p->ran = true;
    return 0;
}

// The derived table copies the plan's contraction reads (split FP32, step
// order), made before a run when the tables changed since.
static int plan_tables(cdet_batch_plan* p) {
    if (p->eval_kind == 3) {
        int rc = ensure_abssum_tables(p->ctx);
        if (rc) return rc;
    }
    if (p->contract_mode == CDET_MODE_SPLIT32) return ensure_split_tables(p->ctx);
    if (p->contract_mode == CDET_MODE_MFMA64) {
        const int rc = ensure_rowsum_tables(p->ctx);
        return rc ? rc : ensure_step_tables(p->ctx);
    }
    return 0;
}

static void plan_free(cdet_batch_plan* p) {
    if (!p) return;
    if (p->ran) (void)hipEventSynchronize(p->done);
    if (p->done) (void)hipEventDestroy(p->done);
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (int k = 0; k + 1 < kMinorLanes; ++k) {
        if (p->aux[k]) (void)hipStreamSynchronize(p->aux[k]);
        if (p->ev_join[k]) (void)hipEventDestroy(p->ev_join[k]);
        (void)hipFree(p->lane_scratch[k + 1].ptr);
    }
    if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    void* bufs[] = {p->d_order, p->d_vbase, p->d_pbase, p->d_obase, p->d_lists, p->d_gup, p->d_gdn, p->d_qkey,
                    p->d_qrank, p->d_qu, p->d_qv, p->d_hist, p->d_keyoff, p->d_tkey, p->d_tstart, p->d_tlen,
                    p->d_ntiles, p->d_su, p->d_sv, p->d_sdest, p->d_out_ord, p->d_single, p->d_fb_list,
                    p->d_fb_count, p->d_wgtot, p->scratch.ptr, p->d_eup, p->d_edn, p->d_prev_verts,
                    p->d_prev_ok, p->d_defer_list, p->d_defer_count};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    delete p;
}

template <typename T>
static int dev_alloc(T** p, size_t n) {
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(1, n) * sizeof(T)));
    return 0;
}

static int contract_launch(cdet_batch_plan* p, hipStream_t s);

// Estimated duration (microseconds) of one order bucket's minors launches on
// an otherwise idle device: a launch floor plus, per round of workgroups, the
// latency of one connected-value recursion (about 3^n dependent terms per
// workgroup; fitted to the order-10 and order-12 buckets of the 4x4 and 6x6
// workloads).
static double bucket_cost_us(int n, int count, int mode, int ncu) {
    const double per_wg = n >= 2 ? 3.3e-4 * std::pow(3.0, n) : 0.0;
    const size_t lds = std::max<size_t>(CdetLds::bytes(n, mode), 1);
    const double fit = std::max<double>(1.0, std::min<double>(8.0, static_cast<double>(kMaxLds / lds)));
    const double rounds = std::max(1.0, 2.0 * count / (static_cast<double>(ncu) * fit));
    return 10.0 + per_wg * rounds;
}

// The plan's minors runs (cdet_batch_plan::runs), highest order first.  The
// mixed run takes the buckets of orders 2 .. kMixedSmallMax and, above, each
// next order's bucket while it holds fewer configurations than the device
// has CUs: such a bucket alone is a chain of short launches that leave the
// device mostly idle, and in the mixed run its workgroups only hold LDS sized
// for the run's highest order.
static void plan_runs(cdet_batch_plan* p) {
    p->runs.clear();
    int lo = 0, hi = 0, nonempty = 0;
    for (int n = 2; n <= kMixedOrderMax; ++n) {
        const int cnt = p->bucket_cnt[static_cast<size_t>(n)];
        if (!cnt) continue;
        if (n > kMixedSmallMax && cnt >= p->ctx->ncu) break;
        if (!lo) lo = n;
        hi = n;
        ++nonempty;
    }
    const bool mix = p->eval_kind != 3 && nonempty >= 2;
    for (int n = kMaxOrder; n >= 1; --n) {
        const int cnt = p->bucket_cnt[static_cast<size_t>(n)];
        if (!cnt) continue;
        if (mix && n >= lo && n <= hi) {
            if (n == hi) {
                const int off = p->bucket_off[static_cast<size_t>(lo)];
                p->runs.push_back({hi, off, p->bucket_off[static_cast<size_t>(hi + 1)] - off, true});
            }
            continue;
        }
        p->runs.push_back({n, p->bucket_off[static_cast<size_t>(n)], cnt, false});
    }
}

// Spread the minors runs over up to max_lanes streams, longest first onto
// the least-loaded lane, when the runs other than the longest carry enough
// work to pay for the fork (about 10-15 us per forked stream); size each
// lane's scratch and create the lanes' streams and events.
static int plan_lanes(cdet_batch_plan* p, hipStream_t* lane_streams) {
    cdet_ctx* ctx = p->ctx;
    p->run_lane.assign(p->runs.size(), 0);
    std::vector<std::pair<double, int>> order;  // (cost, run)
    double total = 0.0, longest = 0.0;
    for (size_t r = 0; r < p->runs.size(); ++r) {
        const auto& run = p->runs[r];
        const double c = bucket_cost_us(run.n, run.cnt, p->eval_kind, ctx->ncu);
        order.emplace_back(c, -static_cast<int>(r));  // ties: the higher order (earlier run) first
        total += c;
        longest = std::max(longest, c);
    }
    std::sort(order.begin(), order.end(), [](const auto& a, const auto& b) {
        return a.first != b.first ? a.first > b.first : a.second > b.second;
    });
    int lanes = 1;
    if (order.size() >= 2 && total - longest >= kForkMinUs)
        lanes = static_cast<int>(std::min<double>({static_cast<double>(p->max_lanes), static_cast<double>(order.size()),
                                                  std::ceil(total / std::max(longest, kForkMinUs))}));
    lanes = std::max(lanes, 1);
    std::vector<double> load(static_cast<size_t>(lanes), 0.0);
    for (const auto& b : order) {
        const int k = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
        load[static_cast<size_t>(k)] += b.first;
        p->run_lane[static_cast<size_t>(-b.second)] = k;
    }
    p->lanes = lanes;
    for (int k = 0; k < lanes; ++k) {
        size_t need = 0;
        for (size_t r = 0; r < p->runs.size(); ++r)
            if (p->run_lane[r] == k)
                need = std::max(need, minors_scratch_bytes(p->runs[r].n, p->runs[r].cnt, p->eval_kind));
        Scratch& sc = k ? p->lane_scratch[k] : p->scratch;
        int rc = grow(&sc.ptr, sc.cap, need);
        if (rc) return rc;
        if (k) {
            // The order lanes stay on ordinary streams: on streams of queues of
            // their own the fork and join of a many-order batch measured
            // slower (3 lanes of a 256-envelope batch: 0.52 against 0.67 M
            // configurations/s).
            if (!lane_streams[k - 1]) HIP_TRY(hipStreamCreateWithFlags(&lane_streams[k - 1], hipStreamNonBlocking));
            p->aux[k - 1] = lane_streams[k - 1];
            HIP_TRY(hipEventCreateWithFlags(&p->ev_join[k - 1], hipEventDisableTiming));
        }
    }
    if (lanes > 1) HIP_TRY(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
    return 0;
}

static int plan_minors(cdet_batch_plan* p, double* d_out, int32_t* d_status, hipStream_t s) {
    cdet_ctx* ctx = p->ctx;
    if (p->lanes > 1) {
        HIP_TRY(hipEventRecord(p->ev_fork, s));
        for (int k = 1; k < p->lanes; ++k) HIP_TRY(hipStreamWaitEvent(p->aux[k - 1], p->ev_fork, 0));
    }
    for (size_t r = 0; r < p->runs.size(); ++r) {
        const auto& run = p->runs[r];
        const int n = run.n;
        const int off = run.off;
        const int lane = p->run_lane[r];
        MinorsOut o{d_out, p->d_out_ord, p->d_single, d_status, p->d_fb_list + 4 * off, p->d_fb_count + n, p->paired};
        o.err_up = p->d_eup;
        o.err_dn = p->d_edn;
        o.defer_list = p->d_defer_list + off;
        o.defer_count = p->d_defer_count + 2 * n;
        int rc = launch_minors(ctx, lane ? p->lane_scratch[lane] : p->scratch, n, p->eval_kind, p->d_lists + off,
                               run.cnt, p->d_pbase, p->d_obase, p->d_gup, p->d_gdn, o, lane ? p->aux[lane - 1] : s,
                               run.mixed ? p->d_order : nullptr,
                               p->contract_mode != CDET_MODE_EXACT && ctx->sum_order == CDET_SUM_TREE);
        if (rc) return rc;
    }
    if (p->lanes > 1) {
        for (int k = 1; k < p->lanes; ++k) {
            HIP_TRY(hipEventRecord(p->ev_join[k - 1], p->aux[k - 1]));
            HIP_TRY(hipStreamWaitEvent(s, p->ev_join[k - 1], 0));
        }
    }
    MinorsOut o{d_out, p->d_out_ord, p->d_single, d_status, nullptr, nullptr, p->paired};
    return launch_combine(p->eval_kind, p->count, p->out_total, p->d_obase, o, s);
}

// Lookup planning and grouping.  The run needs no clearing launches: k_plan
// clears the status words, k_scan_keys clears the histogram after reading it
// and k_cdet clears each bucket's fallback-queue length after use (all start
// cleared at plan creation).
static void plan_group(cdet_batch_plan* p, const cdet_vertex* d_vertices, int32_t* d_status, hipStream_t s) {
    cdet_ctx* ctx = p->ctx;
    // LDS for the row sums of the diagonal pairs (FP64 matrix-core mode): all
    // 2 n rows of the plan's highest order in one pass, up to kPlanRowCapMax
    const int rows_at_v1 = p->contract_mode == CDET_MODE_MFMA64 ? 1 : 0;
    int nmax = 1;
    for (int n = 1; n <= kMaxOrder; ++n)
        if (p->bucket_cnt[static_cast<size_t>(n)]) nmax = n;
    const bool dn_rows = ctx->ks.rowsum[0] != ctx->ks.rowsum[1] || ctx->ks.rowsum[2] != ctx->ks.rowsum[3];
    const int spins = dn_rows ? 2 : 1;
    const int row_cap = rows_at_v1 ? std::min(kPlanRowCapMax, std::max(2 * nmax, 1) * ctx->ks.nu * spins) : 0;
    hipLaunchKernelGGL(k_plan, dim3(static_cast<unsigned>(p->count)), dim3(256),
                       sizeof(double) * static_cast<size_t>(row_cap), s, ctx->ks,
                       reinterpret_cast<const double*>(d_vertices), p->d_order, p->d_vbase, p->d_pbase, p->d_gup,
                       p->d_gdn, p->d_qkey, p->d_qrank, p->d_qu, p->d_qv, p->d_hist, d_status, p->d_eup, p->d_edn,
                       static_cast<double>(ctx->ks.nu + ctx->ks.nv) * kUnitRoundoff, p->reuse, p->d_prev_verts,
                       p->d_prev_ok, ctx->d_pair_cache, rows_at_v1, row_cap);
    const unsigned kwg = static_cast<unsigned>((p->nkeys + kScanKeys - 1) / kScanKeys);
    hipLaunchKernelGGL(k_key_totals, dim3(kwg), dim3(kScanKeys), 0, s, p->d_hist, p->nkeys, p->tile, p->d_wgtot);
    hipLaunchKernelGGL(k_scan_keys, dim3(kwg), dim3(kScanKeys), 0, s, p->d_hist, p->nkeys, p->tile, p->d_wgtot,
                       p->d_keyoff, p->d_tkey, p->d_tstart, p->d_tlen, p->d_ntiles);
    hipLaunchKernelGGL(k_scatter, dim3(static_cast<unsigned>((p->nq + 255) / 256)), dim3(256), 0, s, p->nq, p->d_qkey,
                       p->d_qrank, p->d_qu, p->d_qv, p->d_keyoff, p->d_su, p->d_sv, p->d_sdest);
}

static int plan_launch_sequence(cdet_batch_plan* p, const cdet_vertex* d_vertices, double* d_out, int32_t* d_status,
                                hipStream_t s) {
    plan_group(p, d_vertices, d_status, s);
    int rc = contract_launch(p, s);
    if (rc) return rc;
    rc = plan_minors(p, d_out, d_status, s);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    return 0;
}

// Green-pair reuse: values of an earlier table generation are not reused
// (a fill kernel on the run's stream clears the per-configuration flags).
static int plan_reuse_guard(cdet_batch_plan* p, hipStream_t s) {
    if (!p->reuse || p->reuse_tables == p->ctx->tables_gen) return 0;
    hipLaunchKernelGGL(k_fill_i32, dim3(static_cast<unsigned>((p->count + 255) / 256)), dim3(256), 0, s, p->d_prev_ok,
                       p->count, 0);
    HIP_TRY(hipGetLastError());
    p->reuse_tables = p->ctx->tables_gen;
    return 0;
}

// Capture the plan's launch sequence for these buffers and this stream into
// its graph, unless the graph captured last matches them and the tables.
static int plan_capture(cdet_batch_plan* p, const cdet_vertex* d_vertices, double* d_out, int32_t* d_status,
                        hipStream_t s) {
    if (p->graph_exec && p->g_verts == d_vertices && p->g_out == d_out && p->g_status == d_status && p->g_stream == s &&
        p->g_tables == p->ctx->tables_gen)
        return 0;
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->graph_exec = nullptr;
    p->graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = plan_launch_sequence(p, d_vertices, d_out, d_status, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(CDET_ERR_DEVICE, std::string("graph capture failed: ") + hipGetErrorString(e));
    p->graph = g;
    HIP_TRY(hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0));
    p->g_verts = d_vertices;
    p->g_out = d_out;
    p->g_status = d_status;
    p->g_stream = s;
    p->g_tables = p->ctx->tables_gen;
    return 0;
}

static int contract_launch(cdet_batch_plan* p, hipStream_t s) {
    cdet_ctx* ctx = p->ctx;
    switch (p->contract_mode) {
        case CDET_MODE_EXACT:
            hipLaunchKernelGGL(k_contract_exact, dim3(p->contract_grid), dim3(256), 0, s, ctx->ks, p->d_ntiles,
                               p->d_tkey, p->d_tstart, p->d_tlen, p->d_su, p->d_sv, p->d_sdest, p->d_gup, p->d_gdn);
            break;
        case CDET_MODE_MFMA64:
        case CDET_MODE_SPLIT32:
            return contract_launch_mfma(ctx->ks, p->contract_mode, p->contract_grid, p->d_ntiles, p->d_tkey,
                                        p->d_tstart, p->d_tlen, p->d_su, p->d_sv, p->d_sdest, p->d_gup, p->d_gdn, s,
                                        p->tile);
        default:
            return fail(CDET_ERR_ARGUMENT, "unknown evaluation mode");
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" {

// Plan whose order buckets spread over the run's stream and at most
// max_lanes - 1 of lane_streams (created there on first use).
static int plan_create(cdet_ctx* ctx, const int32_t* orders, int64_t count, int mode, int eval_kind, int max_lanes,
                       hipStream_t* lane_streams, cdet_batch_plan** out) {
    if (!ctx || !orders || !out || count <= 0) return fail(CDET_ERR_ARGUMENT, "cdet_batch_plan_create: bad argument");
    if (mode < 0 || mode > 2) return fail(CDET_ERR_ARGUMENT, "unknown evaluation mode");
    if (!tables_ready(ctx)) return fail(CDET_ERR_TABLES, "Resident Green-function tables have not been uploaded");
    auto* p = new cdet_batch_plan();
    p->ctx = ctx;
    p->max_lanes = std::max(1, std::min({max_lanes, kMinorLanes, ctx->tuning.max_lanes}));
    p->contract_mode = mode;
    if (eval_kind < 0 || eval_kind > 3) {
        delete p;
        return fail(CDET_ERR_ARGUMENT, "unknown evaluation kind");
    }
    p->eval_kind = eval_kind;
    p->count = count;
    std::vector<int64_t> vb(static_cast<size_t>(count)), pb(static_cast<size_t>(count)), ob(static_cast<size_t>(count));
    p->bucket_cnt.assign(kMaxOrder + 1, 0);
    p->bucket_off.assign(kMaxOrder + 2, 0);
    int64_t v = 0, pp = 0, o = 0;
    for (int64_t c = 0; c < count; ++c) {
        const int n = orders[c];
        const int nmin = p->eval_kind == 1 ? 1 : 2;
        if (n < nmin || n > kMaxOrder) {
            const bool order_error = n <= 1 && p->eval_kind != 1;
            plan_free(p);
            if (order_error)
                return fail(CDET_ERR_ORDER, "The all-subset replica integrand is defined only for envelope order >= 2");
            return fail(CDET_ERR_CAPACITY, "configuration order outside the supported range");
        }
        vb[static_cast<size_t>(c)] = v;
        pb[static_cast<size_t>(c)] = pp;
        ob[static_cast<size_t>(c)] = o;
        v += n;
        pp += static_cast<int64_t>(n) * n;
        o += p->eval_kind == 1 ? 1 : (int64_t(1) << n);
        p->bucket_cnt[static_cast<size_t>(n)] += 1;
    }
    p->vert_total = v;
    p->out_total = o;
    p->nq = 2 * pp;
    if (p->nq >= (int64_t(1) << 31)) {
        plan_free(p);
        return fail(CDET_ERR_CAPACITY, "batch too large for one plan");
    }
    for (int n = 1; n <= kMaxOrder; ++n) p->bucket_off[static_cast<size_t>(n + 1)] = p->bucket_off[static_cast<size_t>(n)] + p->bucket_cnt[static_cast<size_t>(n)];
    std::vector<int32_t> lists(static_cast<size_t>(count));
    {
        std::vector<int> fillpos(p->bucket_off.begin(), p->bucket_off.end());
        for (int64_t c = 0; c < count; ++c) lists[static_cast<size_t>(fillpos[static_cast<size_t>(orders[c])]++)] = static_cast<int32_t>(c);
    }
    p->tile = (mode == CDET_MODE_EXACT) ? 64 : mfma_tile_queries(ctx->ks, mode, p->nq, ctx->ncu, ctx->tuning.contract_tile);
    p->nkeys = static_cast<int32_t>(4 * ctx->ks.SP);
    const int64_t nonempty = std::min<int64_t>(p->nkeys, p->nq);
    p->max_tiles = p->nq / p->tile + nonempty + 1;
    int rc = 0;
    rc |= dev_alloc(&p->d_order, count);
    rc |= dev_alloc(&p->d_vbase, count);
    rc |= dev_alloc(&p->d_pbase, count);
    rc |= dev_alloc(&p->d_obase, count);
    rc |= dev_alloc(&p->d_lists, count);
    rc |= dev_alloc(&p->d_gup, p->nq);
    rc |= dev_alloc(&p->d_gdn, p->nq);
    rc |= dev_alloc(&p->d_qkey, p->nq);
    rc |= dev_alloc(&p->d_qrank, p->nq);
    rc |= dev_alloc(&p->d_qu, p->nq);
    rc |= dev_alloc(&p->d_qv, p->nq);
    rc |= dev_alloc(&p->d_hist, static_cast<int64_t>(p->nkeys) * kKeySub);
    rc |= dev_alloc(&p->d_keyoff, static_cast<int64_t>(p->nkeys) * kKeySub);
    rc |= dev_alloc(&p->d_wgtot, (p->nkeys + kScanKeys - 1) / kScanKeys);
    rc |= dev_alloc(&p->d_tkey, p->max_tiles);
    rc |= dev_alloc(&p->d_tstart, p->max_tiles);
    rc |= dev_alloc(&p->d_tlen, p->max_tiles);
    rc |= dev_alloc(&p->d_ntiles, 1);
    rc |= dev_alloc(&p->d_su, p->nq);
    rc |= dev_alloc(&p->d_sv, p->nq);
    rc |= dev_alloc(&p->d_sdest, p->nq);
    rc |= dev_alloc(&p->d_out_ord, p->eval_kind == 1 ? 1 : p->out_total);
    rc |= dev_alloc(&p->d_single, p->eval_kind == 1 ? 4 * count : 1);
    rc |= dev_alloc(&p->d_fb_list, 4 * count);
    rc |= dev_alloc(&p->d_fb_count, kMaxOrder + 1);
    rc |= dev_alloc(&p->d_defer_list, count);
    rc |= dev_alloc(&p->d_defer_count, 2 * (kMaxOrder + 1));
    if (p->eval_kind == 3) {
        rc |= dev_alloc(&p->d_eup, p->nq);
        rc |= dev_alloc(&p->d_edn, p->nq);
    }
    if (!rc && (hipMemset(p->d_hist, 0, sizeof(int32_t) * static_cast<size_t>(p->nkeys) * kKeySub) != hipSuccess ||
                hipMemset(p->d_fb_count, 0, sizeof(int32_t) * (kMaxOrder + 1)) != hipSuccess ||
                hipMemset(p->d_defer_count, 0, sizeof(int32_t) * 2 * (kMaxOrder + 1)) != hipSuccess))
        rc = 1;
    if (!rc && hipEventCreateWithFlags(&p->done, hipEventDisableTiming) != hipSuccess) rc = 1;
    if (rc) {
        plan_free(p);
        return CDET_ERR_DEVICE;
    }
    auto up = [&](void* dst, const void* src, size_t bytes) { return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice); };
    if (up(p->d_order, orders, sizeof(int32_t) * count) != hipSuccess ||
        up(p->d_vbase, vb.data(), sizeof(int64_t) * count) != hipSuccess ||
        up(p->d_pbase, pb.data(), sizeof(int64_t) * count) != hipSuccess ||
        up(p->d_obase, ob.data(), sizeof(int64_t) * count) != hipSuccess ||
        up(p->d_lists, lists.data(), sizeof(int32_t) * count) != hipSuccess) {
        plan_free(p);
        return fail(CDET_ERR_DEVICE, "plan upload failed");
    }
    p->contract_grid = contract_grid(ctx, p->max_tiles, mode, p->tile);
    // Global scratch (minors of every spin matrix, workspaces of large
    // orders), sized now so that runs never allocate and can be graph-captured.
    plan_runs(p);
    {
        int min_cnt = std::numeric_limits<int>::max();
        for (const auto& run : p->runs) min_cnt = std::min(min_cnt, run.cnt);
        p->paired = (p->eval_kind == 0 || p->eval_kind == 2) &&
                    static_cast<int64_t>(min_cnt) >= static_cast<int64_t>(ctx->tuning.pair_min_per_cu) * ctx->ncu;
    }
    rc = plan_lanes(p, lane_streams);
    if (rc) {
        plan_free(p);
        return rc;
    }
    rc = plan_tables(p);
    if (rc) {
        plan_free(p);
        return rc;
    }
    *out = p;
    return 0;
}

int cdet_batch_plan_create(cdet_ctx* ctx, const int32_t* orders, int64_t count, int mode, int eval_kind,
                           cdet_batch_plan** out) {
    if (!ctx) return fail(CDET_ERR_ARGUMENT, "cdet_batch_plan_create: bad argument");
    return plan_create(ctx, orders, count, mode, eval_kind, kMinorLanes, ctx->lanes, out);
}

int cdet_batch_plan_destroy(cdet_batch_plan* plan) {
    plan_free(plan);
    return 0;
}

int cdet_batch_plan_reuse_pairs(cdet_batch_plan* p, int enable) {
    if (!p) return fail(CDET_ERR_ARGUMENT, "cdet_batch_plan_reuse_pairs: null plan");
    enable = enable ? 1 : 0;
    if (p->eval_kind == 3) enable = 0;  // the error bounds need every pair's own error scale
    if (enable == p->reuse) return 0;
    if (p->ran) (void)hipEventSynchronize(p->done);
    if (enable && !p->d_prev_verts) {
        int rc = dev_alloc(&p->d_prev_verts, 4 * p->vert_total);
        rc |= dev_alloc(&p->d_prev_ok, p->count);
        if (rc) return CDET_ERR_DEVICE;
    }
    if (enable) HIP_TRY(hipMemset(p->d_prev_ok, 0, sizeof(int32_t) * static_cast<size_t>(p->count)));
    p->reuse = enable;
    p->reuse_tables = p->ctx->tables_gen;
    // the captured graph holds the planning kernel's arguments: capture again
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->graph_exec = nullptr;
    p->graph = nullptr;
    return 0;
}

int64_t cdet_batch_out_size(const cdet_batch_plan* plan) { return plan ? plan->out_total : -1; }
int64_t cdet_batch_vertex_count(const cdet_batch_plan* plan) { return plan ? plan->vert_total : -1; }
int64_t cdet_batch_query_count(const cdet_batch_plan* plan) { return plan ? plan->nq : -1; }
int cdet_batch_plan_streams(const cdet_batch_plan* plan) { return plan ? plan->lanes : -1; }

int cdet_batch_contracted(cdet_batch_plan* p, int64_t* lookups, int64_t* tiles) {
    if (!p || !lookups || !tiles) return fail(CDET_ERR_ARGUMENT, "cdet_batch_contracted: bad argument");
    if (!p->ran) return fail(CDET_ERR_ARGUMENT, "cdet_batch_contracted: the plan has not run");
    HIP_TRY(hipEventSynchronize(p->done));
    int32_t nt = 0;
    HIP_TRY(hipMemcpy(&nt, p->d_ntiles, sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int32_t> len(static_cast<size_t>(std::max(nt, 0)));
    if (nt > 0) HIP_TRY(hipMemcpy(len.data(), p->d_tlen, sizeof(int32_t) * len.size(), hipMemcpyDeviceToHost));
    int64_t total = 0;
    for (int32_t l : len) total += l;
    *lookups = total;
    *tiles = nt;
    return 0;
}

int cdet_batch_run(cdet_batch_plan* p, const cdet_vertex* d_vertices, double* d_out, int32_t* d_status,
                   int use_graph, void* stream) {
    if (!p || !d_vertices || !d_out) return fail(CDET_ERR_ARGUMENT, "cdet_batch_run: bad argument");
    int trc = plan_tables(p);
    if (trc) return trc;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : p->ctx->stream;
    trc = plan_reuse_guard(p, s);
    if (trc) return trc;
    if (!use_graph) {
        int rc = plan_launch_sequence(p, d_vertices, d_out, d_status, s);
        if (!rc) rc = plan_mark_done(p, s);
        if (!rc) p->ctx->calls += static_cast<uint64_t>(p->count);
        return rc;
    }
    int rc = plan_capture(p, d_vertices, d_out, d_status, s);
    if (rc) return rc;
    HIP_TRY(hipGraphLaunch(p->graph_exec, s));
    rc = plan_mark_done(p, s);
    if (rc) return rc;
    p->ctx->calls += static_cast<uint64_t>(p->count);
    return 0;
}

int cdet_batch_profile(cdet_batch_plan* p, const cdet_vertex* d_vertices, double* d_out, int32_t* d_status,
                       void* stream, float* stage_ms) {
    if (!p || !d_vertices || !d_out || !stage_ms) return fail(CDET_ERR_ARGUMENT, "cdet_batch_profile: bad argument");
    int trc = plan_tables(p);
    if (trc) return trc;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : p->ctx->stream;
    trc = plan_reuse_guard(p, s);
    if (trc) return trc;
    hipEvent_t ev[4];
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipEventRecord(ev[0], s));
    plan_group(p, d_vertices, d_status, s);
    HIP_TRY(hipEventRecord(ev[1], s));
    int rc = 0;
    for (int r = 0; r < CDET_PROFILE_CONTRACT_REPS && !rc; ++r) rc = contract_launch(p, s);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(ev[2], s));
    rc = plan_minors(p, d_out, d_status, s);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(ev[3], s));
    rc = plan_mark_done(p, s);
    if (rc) return rc;
    HIP_TRY(hipEventSynchronize(ev[3]));
    for (int k = 0; k < 3; ++k) HIP_TRY(hipEventElapsedTime(&stage_ms[k], ev[k], ev[k + 1]));
    stage_ms[1] /= CDET_PROFILE_CONTRACT_REPS;
    for (auto& e : ev) (void)hipEventDestroy(e);
    return 0;
}

int cdet_eval_batch(cdet_ctx* ctx, const cdet_vertex* d_vertices, const int32_t* h_orders, int64_t count,
                    double* d_out, int32_t* d_status, int mode, void* stream) {
    cdet_batch_plan* p = nullptr;
    int rc = cdet_batch_plan_create(ctx, h_orders, count, mode, 0, &p);
    if (rc) return rc;
    rc = cdet_batch_run(p, d_vertices, d_out, d_status, 0, stream);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    (void)hipStreamSynchronize(s);
    plan_free(p);
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Table evaluation only (cdet_green_pair_eval): lookups grouped by coefficient
// block and contracted by the same kernels as the integrand batches.
// ---------------------------------------------------------------------------
// ev non-null: events ev[0..3] around planning and grouping, the contraction
// (CDET_PROFILE_CONTRACT_REPS launches back to back) and the interleave.
static int green_pair_run(cdet_ctx* ctx, const int32_t* query_ints, const double* query_taus, int64_t count,
                          double* out, int mode, void* stream, hipEvent_t* ev) {
    if (!ctx || !query_ints || !query_taus || !out || count <= 0)
        return fail(CDET_ERR_ARGUMENT, "cdet_green_pair_eval: bad argument");
    if (!tables_ready(ctx)) return fail(CDET_ERR_TABLES, "Resident Green-function tables have not been uploaded");
    if (count >= (int64_t(1) << 31)) return fail(CDET_ERR_CAPACITY, "too many lookups for one call");
    if (mode == CDET_MODE_SPLIT32 || mode == CDET_MODE_MFMA64) {
        int rc = mode == CDET_MODE_SPLIT32 ? ensure_split_tables(ctx) : ensure_step_tables(ctx);
        if (rc) return rc;
    }
    // The context's lookup scratch is shared by every call: one call at a
    // time sets it up and enqueues (query_mu), and each call's launches wait
    // on the device for the previous call's to finish.
    std::lock_guard<std::mutex> lock(ctx->query_mu);
    QueryScratch& q = ctx->query;
    const int32_t nkeys = static_cast<int32_t>(4 * ctx->ks.SP);
    const int tile =
        mode == CDET_MODE_EXACT ? 64 : mfma_tile_queries(ctx->ks, mode, count, ctx->ncu, ctx->tuning.contract_tile);
    const int64_t max_tiles = count / tile + std::min<int64_t>(nkeys, count) + 1;
    if (q.cap < count || q.tiles_cap < max_tiles || q.nkeys != nkeys || !q.hist) {
        query_scratch_free(q);
        int rc = 0;
        rc |= dev_alloc(&q.gup, count);
        rc |= dev_alloc(&q.gdn, count);
        rc |= dev_alloc(&q.qu, count);
        rc |= dev_alloc(&q.qv, count);
        rc |= dev_alloc(&q.su, count);
        rc |= dev_alloc(&q.sv, count);
        rc |= dev_alloc(&q.qkey, count);
        rc |= dev_alloc(&q.qrank, count);
        rc |= dev_alloc(&q.sdest, count);
        rc |= dev_alloc(&q.hist, static_cast<int64_t>(nkeys) * kKeySub);
        rc |= dev_alloc(&q.keyoff, static_cast<int64_t>(nkeys) * kKeySub);
        rc |= dev_alloc(&q.wgtot, (nkeys + kScanKeys - 1) / kScanKeys);
        rc |= dev_alloc(&q.tkey, max_tiles);
        rc |= dev_alloc(&q.tstart, max_tiles);
        rc |= dev_alloc(&q.tlen, max_tiles);
        rc |= dev_alloc(&q.ntiles, 1);
        if (rc) {
            query_scratch_free(q);
            return CDET_ERR_DEVICE;
        }
        q.cap = count;
        q.tiles_cap = max_tiles;
        q.nkeys = nkeys;
    }
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    HIP_TRY(hipStreamWaitEvent(s, q.done, 0));
    if (ev) HIP_TRY(hipEventRecord(ev[0], s));
    const unsigned gq = static_cast<unsigned>((count + 255) / 256);
    hipLaunchKernelGGL(k_fill_i32, dim3(static_cast<unsigned>((nkeys * kKeySub + 255) / 256)), dim3(256), 0, s,
                       q.hist, static_cast<int64_t>(nkeys) * kKeySub, 0);
    hipLaunchKernelGGL(k_plan_queries, dim3(gq), dim3(256), 0, s, ctx->ks, count, query_ints, query_taus, q.gup, q.gdn,
                       q.qkey, q.qrank, q.qu, q.qv, q.hist);
    const unsigned kwg = static_cast<unsigned>((nkeys + kScanKeys - 1) / kScanKeys);
    hipLaunchKernelGGL(k_key_totals, dim3(kwg), dim3(kScanKeys), 0, s, q.hist, nkeys, tile, q.wgtot);
    hipLaunchKernelGGL(k_scan_keys, dim3(kwg), dim3(kScanKeys), 0, s, q.hist, nkeys, tile, q.wgtot, q.keyoff, q.tkey,
                       q.tstart, q.tlen, q.ntiles);
    hipLaunchKernelGGL(k_scatter, dim3(gq), dim3(256), 0, s, count, q.qkey, q.qrank, q.qu, q.qv, q.keyoff, q.su, q.sv,
                       q.sdest);
    const int grid = contract_grid(ctx, max_tiles, mode, tile);
    if (ev) HIP_TRY(hipEventRecord(ev[1], s));
    for (int r = 0; r < (ev ? CDET_PROFILE_CONTRACT_REPS : 1); ++r) {
        if (mode == CDET_MODE_EXACT) {
            hipLaunchKernelGGL(k_contract_exact, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, ctx->ks,
                               q.ntiles, q.tkey, q.tstart, q.tlen, q.su, q.sv, q.sdest, q.gup, q.gdn);
        } else {
            int rc = contract_launch_mfma(ctx->ks, mode, static_cast<int>(grid), q.ntiles, q.tkey, q.tstart, q.tlen,
                                          q.su, q.sv, q.sdest, q.gup, q.gdn, s, tile);
            if (rc) return rc;
        }
    }
    if (ev) HIP_TRY(hipEventRecord(ev[2], s));
    hipLaunchKernelGGL(k_interleave, dim3(gq), dim3(256), 0, s, count, q.gup, q.gdn, out);
    HIP_TRY(hipGetLastError());
    if (ev) HIP_TRY(hipEventRecord(ev[3], s));
    HIP_TRY(hipEventRecord(q.done, s));
    return 0;
}

extern "C" int cdet_green_pair_eval(cdet_ctx* ctx, const int32_t* query_ints, const double* query_taus, int64_t count,
                                    double* out, int mode, void* stream) {
    return green_pair_run(ctx, query_ints, query_taus, count, out, mode, stream, nullptr);
}

extern "C" int cdet_green_pair_profile(cdet_ctx* ctx, const int32_t* query_ints, const double* query_taus,
                                       int64_t count, double* out, int mode, void* stream, float* stage_ms) {
    if (!stage_ms) return fail(CDET_ERR_ARGUMENT, "cdet_green_pair_profile: bad argument");
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int rc = 0;
    for (auto& e : ev)
        if (!rc && hipEventCreate(&e) != hipSuccess) rc = fail(CDET_ERR_DEVICE, "hipEventCreate failed");
    if (!rc) rc = green_pair_run(ctx, query_ints, query_taus, count, out, mode, stream, ev);
    if (!rc && hipEventSynchronize(ev[3]) != hipSuccess) rc = fail(CDET_ERR_DEVICE, "hipEventSynchronize failed");
    for (int k = 0; !rc && k < 3; ++k)
        if (hipEventElapsedTime(&stage_ms[k], ev[k], ev[k + 1]) != hipSuccess)
            rc = fail(CDET_ERR_DEVICE, "hipEventElapsedTime failed");
    if (!rc) stage_ms[1] /= CDET_PROFILE_CONTRACT_REPS;
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

extern "C" int cdet_tables_build(cdet_ctx* ctx, const cdet_lattice* lattice) {
    return build_tables_gpu(ctx, lattice);
}

// ---------------------------------------------------------------------------
// Host-buffer batch evaluation for the driver (cdet_batch_eval_fn).
// ---------------------------------------------------------------------------
extern "C" int cdet_ctx_set_pair_reuse(cdet_ctx* ctx, int enable) {
    if (!ctx) return fail(CDET_ERR_ARGUMENT, "cdet_ctx_set_pair_reuse: null context");
    ctx->pair_reuse = enable ? 1 : 0;
    if (ctx->host_plan) {
        int rc = cdet_batch_plan_reuse_pairs(ctx->host_plan, ctx->pair_reuse);
        if (rc) return rc;
    }
    for (auto& sl : ctx->slots) {
        if (!sl.plan) continue;
        if (sl.pending) return fail(CDET_ERR_ARGUMENT, "cdet_ctx_set_pair_reuse: a submitted batch is pending");
        int rc = cdet_batch_plan_reuse_pairs(sl.plan, ctx->pair_reuse);
        if (rc) return rc;
    }
    return 0;
}

extern "C" int cdet_ctx_set_batch_mode(cdet_ctx* ctx, int mode) {
    if (!ctx || mode < 0 || mode > 2) return fail(CDET_ERR_ARGUMENT, "cdet_ctx_set_batch_mode: bad argument");
    if (ctx->batch_mode != mode && ctx->host_plan) {
        cdet_batch_plan_destroy(ctx->host_plan);
        ctx->host_plan = nullptr;
        ctx->host_plan_orders.clear();
    }
    if (ctx->batch_mode != mode) {
        for (auto& sl : ctx->slots) {
            if (sl.pending) return fail(CDET_ERR_ARGUMENT, "cdet_ctx_set_batch_mode: a submitted batch is pending");
            if (sl.plan) cdet_batch_plan_destroy(sl.plan);
            sl.plan = nullptr;
            sl.orders.clear();
        }
    }
    ctx->batch_mode = mode;
    return 0;
}

// Pinned host staging that only grows.
static int grow_pinned(void** p, size_t& cap, size_t need) {
    if (need <= cap) return 0;
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    cap = 0;
    HIP_TRY(hipHostMalloc(p, need, hipHostMallocDefault));
    cap = need;
    return 0;
}

extern "C" void* cdet_host_alloc(void* user, size_t bytes) {
    cdet_ctx* ctx = static_cast<cdet_ctx*>(user);
    if (!ctx || bytes == 0) {
        fail(CDET_ERR_ARGUMENT, "cdet_host_alloc: bad argument");
        return nullptr;
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        fail(CDET_ERR_DEVICE, "cdet_host_alloc: hipHostMalloc failed");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(ctx->host_mu);
    ctx->host_bufs[static_cast<const char*>(p)] = bytes;
    return p;
}

extern "C" void cdet_host_free(void* user, void* p) {
    cdet_ctx* ctx = static_cast<cdet_ctx*>(user);
    if (!ctx || !p) return;
    std::lock_guard<std::mutex> lock(ctx->host_mu);
    auto it = ctx->host_bufs.find(static_cast<const char*>(p));
    if (it == ctx->host_bufs.end()) return;
    ctx->host_bufs.erase(it);
    (void)hipHostFree(p);
}

// [p, p + bytes) inside one of the context's cdet_host_alloc buffers (host_mu held).
static bool in_host_buf(const cdet_ctx* ctx, const void* p, size_t bytes) {
    const char* c = static_cast<const char*>(p);
    auto it = ctx->host_bufs.upper_bound(c);
    if (it == ctx->host_bufs.begin()) return false;
    --it;
    return c >= it->first && c + bytes <= it->first + it->second;
}

extern "C" int cdet_eval_batch_host(void* user, const cdet_vertex* h_vertices, const int32_t* h_orders, int64_t count,
                                    double* h_out) {
    cdet_ctx* ctx = static_cast<cdet_ctx*>(user);
    if (!ctx || !h_vertices || !h_orders || !h_out || count <= 0)
        return fail(CDET_ERR_ARGUMENT, "cdet_eval_batch_host: bad argument");
    std::lock_guard<std::mutex> lock(ctx->host_mu);
    const bool same = ctx->host_plan && static_cast<int64_t>(ctx->host_plan_orders.size()) == count &&
                      std::equal(ctx->host_plan_orders.begin(), ctx->host_plan_orders.end(), h_orders);
    if (!same) {
        if (ctx->host_plan) cdet_batch_plan_destroy(ctx->host_plan);
        ctx->host_plan = nullptr;
        int rc = cdet_batch_plan_create(ctx, h_orders, count, ctx->batch_mode, 0, &ctx->host_plan);
        if (rc) return rc;
        // consecutive calls are steps of the same chains: reuse the Green
        // pairs of retained vertices, as the reference's pair cache does
        rc = cdet_batch_plan_reuse_pairs(ctx->host_plan, ctx->pair_reuse);
        if (rc) return rc;
        ctx->host_plan_orders.assign(h_orders, h_orders + count);
    }
    cdet_batch_plan* p = ctx->host_plan;
    const size_t vbytes = sizeof(cdet_vertex) * static_cast<size_t>(p->vert_total);
    const size_t obytes = sizeof(double) * static_cast<size_t>(p->out_total);
    const size_t sbytes = sizeof(int32_t) * static_cast<size_t>(count);
    int rc = grow(&ctx->d_hv, ctx->d_hv_cap, vbytes);
    if (!rc) rc = grow(&ctx->d_hout, ctx->d_hout_cap, obytes);
    if (!rc) rc = grow(&ctx->d_hstatus, ctx->d_hstatus_cap, sbytes);
    // pinned staging: the copies run as DMA transfers instead of going
    // through the runtime's bounce buffer for pageable memory
    if (!rc) rc = grow_pinned(&ctx->h_hv, ctx->h_hv_cap, vbytes);
    if (!rc) rc = grow_pinned(&ctx->h_hout, ctx->h_hout_cap, obytes);
    if (!rc) rc = grow_pinned(&ctx->h_hstatus, ctx->h_hstatus_cap, sbytes);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    std::memcpy(ctx->h_hv, h_vertices, vbytes);
    HIP_TRY(hipMemcpyAsync(ctx->d_hv, ctx->h_hv, vbytes, hipMemcpyHostToDevice, s));
    rc = cdet_batch_run(p, static_cast<const cdet_vertex*>(ctx->d_hv), static_cast<double*>(ctx->d_hout),
                        static_cast<int32_t*>(ctx->d_hstatus), 1, s);
    if (rc) return rc;
    // values straight into a pinned caller buffer, else through the staging
    const bool direct = in_host_buf(ctx, h_out, obytes);
    HIP_TRY(hipMemcpyAsync(direct ? static_cast<void*>(h_out) : ctx->h_hout, ctx->d_hout, obytes,
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(ctx->h_hstatus, ctx->d_hstatus, sbytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const int32_t* st = static_cast<const int32_t*>(ctx->h_hstatus);
    for (int64_t i = 0; i < count; ++i)
        if (st[i] & 1) return fail(CDET_ERR_NONFINITE, "Error (all-subset integrand): Value is not finite");
    if (!direct) std::memcpy(h_out, ctx->h_hout, obytes);
    return 0;
}

// The slot's launch: vertices to the device, the plan's graph (captured by
// the submitting thread) on the slot's stream, values and status words back
// to pinned memory.
static int slot_launch(cdet_ctx::HostSlot& sl) {
    cdet_batch_plan* p = sl.plan;
    HIP_TRY(hipMemcpyAsync(sl.d_v, sl.h_v, sizeof(cdet_vertex) * static_cast<size_t>(p->vert_total),
                           hipMemcpyHostToDevice, sl.stream));
    HIP_TRY(hipGraphLaunch(p->graph_exec, sl.stream));
    int rc = plan_mark_done(p, sl.stream);
    if (rc) return rc;
    p->ctx->calls += static_cast<uint64_t>(p->count);
    HIP_TRY(hipMemcpyAsync(sl.h_out, sl.d_out, sl.out_bytes, hipMemcpyDeviceToHost, sl.stream));
    HIP_TRY(hipMemcpyAsync(sl.h_status, sl.d_status, sizeof(int32_t) * static_cast<size_t>(sl.count),
                           hipMemcpyDeviceToHost, sl.stream));
    return 0;
}

static void slot_worker(int device, cdet_ctx::HostSlot* sl) {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(sl->mu);
    for (;;) {
        sl->cv.wait(lk, [sl] { return sl->quit || sl->job; });
        if (sl->quit) return;
        sl->job = false;
        lk.unlock();
        const int rc = slot_launch(*sl);
        std::string err = rc ? g_last_error : std::string();
        lk.lock();
        sl->launch_rc = rc;
        sl->launch_err = std::move(err);
        sl->launched = true;
        sl->cv.notify_all();
    }
}

// Asynchronous form of cdet_eval_batch_host over two slots: the vertices are
// copied into the slot's pinned staging and the slot's thread launches the
// copies and the plan on the slot's stream; cdet_eval_batch_wait waits for
// the launch and the stream and copies the values out.
extern "C" int cdet_eval_batch_submit(void* user, int slot, const cdet_vertex* h_vertices, const int32_t* h_orders,
                                      int64_t count) {
    cdet_ctx* ctx = static_cast<cdet_ctx*>(user);
    if (!ctx || slot < 0 || slot > 1 || !h_vertices || !h_orders || count <= 0)
        return fail(CDET_ERR_ARGUMENT, "cdet_eval_batch_submit: bad argument");
    auto& sl = ctx->slots[slot];
    if (sl.pending) return fail(CDET_ERR_ARGUMENT, "cdet_eval_batch_submit: the slot's previous batch was not waited for");
    if (!sl.stream) HIP_TRY(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
    const bool same = sl.plan && static_cast<int64_t>(sl.orders.size()) == count &&
                      std::equal(sl.orders.begin(), sl.orders.end(), h_orders);
    if (!same) {
        if (sl.plan) {
            (void)hipStreamSynchronize(sl.stream);
            cdet_batch_plan_destroy(sl.plan);
        }
        sl.plan = nullptr;
        int rc = plan_create(ctx, h_orders, count, ctx->batch_mode, 0, kSlotLanes, sl.lanes, &sl.plan);
        if (rc) return rc;
        rc = cdet_batch_plan_reuse_pairs(sl.plan, ctx->pair_reuse);
        if (rc) return rc;
        sl.orders.assign(h_orders, h_orders + count);
    }
    cdet_batch_plan* p = sl.plan;
    // derived table copies are made here, so that the slot threads only read
    // the context's tables
    int rc = plan_tables(p);
    if (rc) return rc;
    const size_t vbytes = sizeof(cdet_vertex) * static_cast<size_t>(p->vert_total);
    const size_t obytes = sizeof(double) * static_cast<size_t>(p->out_total);
    const size_t sbytes = sizeof(int32_t) * static_cast<size_t>(count);
    rc = grow(&sl.d_v, sl.d_v_cap, vbytes);
    if (!rc) rc = grow(&sl.d_out, sl.d_out_cap, obytes);
    if (!rc) rc = grow(&sl.d_status, sl.d_status_cap, sbytes);
    if (!rc) rc = grow_pinned(&sl.h_v, sl.h_v_cap, vbytes);
    if (!rc) rc = grow_pinned(&sl.h_out, sl.h_out_cap, obytes);
    if (!rc) rc = grow_pinned(&sl.h_status, sl.h_status_cap, sbytes);
    if (rc) return rc;
    rc = plan_reuse_guard(p, sl.stream);
    if (rc) return rc;
    // the graph is captured here, on the submitting thread, so that the slot
    // threads make only stream-ordered calls (copies, graph launches)
    rc = plan_capture(p, static_cast<const cdet_vertex*>(sl.d_v), static_cast<double*>(sl.d_out),
                      static_cast<int32_t*>(sl.d_status), sl.stream);
    if (rc) return rc;
    std::memcpy(sl.h_v, h_vertices, vbytes);
    sl.count = count;
    sl.out_bytes = obytes;
    if (!sl.worker.joinable()) sl.worker = std::thread(slot_worker, ctx->device, &sl);
    {
        std::lock_guard<std::mutex> lk(sl.mu);
        sl.job = true;
        sl.launched = false;
    }
    sl.cv.notify_all();
    sl.pending = true;
    return 0;
}

extern "C" int cdet_eval_batch_wait(void* user, int slot, double* h_out) {
    cdet_ctx* ctx = static_cast<cdet_ctx*>(user);
    if (!ctx || slot < 0 || slot > 1 || !h_out) return fail(CDET_ERR_ARGUMENT, "cdet_eval_batch_wait: bad argument");
    auto& sl = ctx->slots[slot];
    if (!sl.pending) return fail(CDET_ERR_ARGUMENT, "cdet_eval_batch_wait: no batch was submitted to the slot");
    {
        std::unique_lock<std::mutex> lk(sl.mu);
        sl.cv.wait(lk, [&sl] { return sl.launched; });
    }
    sl.pending = false;
    if (sl.launch_rc) {
        // copies queued before the failure may still read the staging
        (void)hipStreamSynchronize(sl.stream);
        return fail(sl.launch_rc, sl.launch_err);
    }
    HIP_TRY(hipStreamSynchronize(sl.stream));
    const int32_t* st = static_cast<const int32_t*>(sl.h_status);
    for (int64_t i = 0; i < sl.count; ++i)
        if (st[i] & 1) return fail(CDET_ERR_NONFINITE, "Error (all-subset integrand): Value is not finite");
    std::memcpy(h_out, sl.h_out, sl.out_bytes);
    return 0;
}


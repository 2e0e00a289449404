"""GPU parity at the five benchmark workloads, against the oracle.

C1 (3x3, 64 coefficients) is covered by test_gpu_parity.py and
test_gpu_driver.py.  Here:

* C2 (4x4, beta 6, 2x2 corner block, 96 coefficients): oracle-built tables; a
  slice of the seeded 8192-configuration set covering every order 1..10, with
  64 order-10 configurations, evaluated as all-subset and single-order
  integrands by the batched FP64 form and the plain form: bit-identical.
* C3 (8x8, beta 8, 128 coefficients): GPU-built tables; the 2^20 seeded
  twisted-table lookups, a sample checked against the oracle's Clenshaw on
  the same coefficient blocks (FP64 form bit-identical, FP64 matrix-core form
  within its summation bound), and the table against the direct Green
  function at the sampled times.
* C4 (6x6, beta 8, x < 3 half strip, 128 coefficients): GPU-built tables; the
  256 seeded chain envelopes (orders 2..12), a sample across every order
  checked against the oracle: bit-identical.
* C5 (10x10, beta 10, x < 5 half, 192 coefficients): tables built on the GPU in
  uneven node chunks (7 Nc + 3); the 1024 seeded configurations (orders
  2..14), a sample with orders 13 and 14 checked against the oracle
  bit-identical, all-subset[full] against single-order at orders 13-14, and
  the 192-coefficient table against the direct Green function.

The oracle reads the large GPU-built tables through helpers.sparse_tables:
only the coefficient blocks the checked lookups read are copied back.
"""
import ctypes
import sys

import numpy as np
import pytest

import bench
import cdet_amd
import helpers
import rng
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits_equal(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def oracle_params(w, alpha_key="alpha"):
    return O.make_params(w["N1"], w["N2"], w["nc"], w["beta"], w["U"], w["mu"], w[alpha_key], w[alpha_key])


def gpu_context(w, op):
    p = cdet_amd.make_params(w["N1"], w["N2"], w["nc"], w["beta"], w["U"], w["mu"], w["alpha"], w["alpha"],
                             derived=(op.scale, op.diag_up_shift, op.diag_down_shift))
    return cdet_amd.Context(p)


def run_plan(ctx, configs, mode, single=False, kind=None):
    import torch
    orders = [len(v) for v in configs]
    plan = ctx.plan(orders, mode=mode, single=single, kind=kind)
    verts = torch.tensor(np.concatenate(configs), dtype=torch.float64, device="cuda")
    out = torch.full((plan.out_size,), float("nan"), dtype=torch.float64, device="cuda")
    status = torch.zeros(len(orders), dtype=torch.int32, device="cuda")
    plan.run(verts, out, status, use_graph=False, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    plan.close()
    _, ob = cdet_amd.packed_offsets(orders, single)
    return out.cpu().numpy(), ob, status.cpu().numpy()


def sparse_oracle(ctx, op, configs):
    """Oracle over the device tables' blocks that `configs` read (equal spin
    shifts: kinds 1 and 3 alias 0 and 2)."""
    p = ctx.params
    tw, od = set(), set()
    for v in configs:
        a, b = helpers.config_sitepairs(v, p.N1, p.N2)
        tw |= a
        od |= b
    tables = helpers.sparse_tables(ctx, [tw, tw, od, od])
    return O.Oracle(op, tables)


def pick_per_order(orders, per_order, extra=None):
    """Indices: the first `per_order` configurations of every order present,
    plus the first extra[n] of order n."""
    sel = []
    for n in sorted(set(int(x) for x in orders)):
        k = per_order if extra is None or n not in extra else max(per_order, extra[n])
        sel += list(np.nonzero(orders == n)[0][:k])
    return sorted(int(i) for i in sel)


# ---------------------------------------------------------------------------
# C2
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2():
    w = bench.WORKLOAD
    say("c2: oracle tables (4x4, 96 coefficients)")
    tables = O.build_tables(w["N1"], w["N2"], w["beta"], w["mu"], w["alpha"], w["alpha"], w["nc"],
                            mask=w["mask"], mask_param=w["radius"])
    op = oracle_params(w)
    ctx = gpu_context(w, op)
    ctx.upload_tables(tables)
    orders, verts = rng.synthetic_configs(w["base_seed"], w["tag"], w["configs"])
    yield dict(w=w, ctx=ctx, orc=O.Oracle(op, tables), orders=orders, verts=verts, op=op, tables=tables)
    ctx.close()


def test_c2_batched_exact_bit_identical(gpu_available, c2):
    orders, verts, orc = c2["orders"], c2["verts"], c2["orc"]
    idx = pick_per_order(orders, 6, extra={10: 64})
    assert sum(1 for i in idx if orders[i] == 10) == 64
    sub = [i for i in idx if orders[i] >= 2]
    configs = [verts[i] for i in sub]
    say(f"c2: {len(configs)} all-subset configurations (orders 2..10), exact")
    out, ob, st = run_plan(c2["ctx"], configs, "exact")
    assert not np.any(st & 1)
    for k, v in enumerate(configs):
        assert bits_equal(out[ob[k]:ob[k + 1]], orc.all_subsets(v)), (sub[k], len(v))
    configs1 = [verts[i] for i in idx]
    say(f"c2: {len(configs1)} single-order configurations (orders 1..10), exact")
    out1, _, _ = run_plan(c2["ctx"], configs1, "exact", single=True)
    for k, v in enumerate(configs1):
        assert bits_equal(out1[k], orc.single(v)), (idx[k], len(v))


def test_c2_plain_form_bit_identical(gpu_available, c2):
    orders, verts, orc = c2["orders"], c2["verts"], c2["orc"]
    for i in list(np.nonzero(orders == 10)[0][:3]) + list(np.nonzero(orders == 7)[0][:2]):
        say(f"c2: plain form configuration {i} (order {orders[i]})")
        got = c2["ctx"].evaluate_all_subset_integrands(verts[i])
        assert bits_equal(got, orc.all_subsets(verts[i])), i
        assert bits_equal(c2["ctx"].evaluate_integrand(verts[i]), orc.single(verts[i])), i


def test_c2_constant_tables_exact_rationals_on_gpu(gpu_available):
    # the exact-rational pin of the assembly (prefactor, singletons,
    # twisted - 2 ordinary) through the GPU path: plain and batched
    from test_oracle_integrand import EXACT, constant_tables, exact_values
    from fractions import Fraction
    w = EXACT
    p = cdet_amd.make_params(w["N1"], w["N2"], w["nc"], w["beta"], w["U"], w["mu"], w["alpha_up"], w["alpha_down"])
    assert p.scale == 2.0 and p.diag_up_shift == 0.25 and p.diag_down_shift == 0.5
    ctx = cdet_amd.Context(p)
    ctx.upload_tables(constant_tables(0.25, -0.125))
    n = 9
    want, scale = exact_values(n, 0.25, -0.125)
    v = helpers.seeded_configs(4242, [n])[0]
    say("exact rationals on the GPU: plain and batched")
    got_plain = ctx.evaluate_all_subset_integrands(v)
    got_batch, _, _ = run_plan(ctx, [v], "exact")
    for got in (got_plain, got_batch):
        for m in range(1, 1 << n):
            k = bin(m).count("1")
            assert abs(Fraction(got[m]) - want[k]) <= Fraction(1e-13) * scale[k], (m, k)
    for k in range(1, 9):
        vk = helpers.seeded_configs(4343 + k, [k])[0]
        assert abs(Fraction(ctx.evaluate_integrand(vk)) - want[k]) <= Fraction(1e-13) * scale[k], k
    ctx.close()


# ---------------------------------------------------------------------------
# C3
# ---------------------------------------------------------------------------
def green_pair_eval(ctx, d_qi, d_qt, n, mode):
    import torch
    out = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
    rc = cdet_amd.lib().cdet_green_pair_eval(ctx.handle, ctypes.c_void_p(d_qi.data_ptr()),
                                             ctypes.c_void_p(d_qt.data_ptr()), n, ctypes.c_void_p(out.data_ptr()),
                                             cdet_amd.MODES[mode], None)
    assert rc == 0, cdet_amd._err()
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(n, 2)


def test_c3_lookups(gpu_available):
    import torch
    w = bench.C3
    op = oracle_params(w)
    ctx = gpu_context(w, op)
    say("c3: building tables (8x8, 128 coefficients)")
    ctx.build_tables(pbc=True, mask=w["mask"], mask_param=w["radius"])
    qi, qt = bench.c3_queries()
    n = len(qi)
    assert n == 1 << 20
    d_qi = torch.tensor(qi, device="cuda")
    d_qt = torch.tensor(qt, device="cuda")
    say("c3: 2^20 lookups, exact and mfma64")
    ex = green_pair_eval(ctx, d_qi, d_qt, n, "exact")
    mf = green_pair_eval(ctx, d_qi, d_qt, n, "mfma64")
    r = np.random.default_rng(33)
    samp = np.sort(r.choice(n, 2048, replace=False))
    N1, N2 = w["N1"], w["N2"]
    sps = {helpers.sitepair_index(*qi[k, 1:7], N1, N2) for k in samp}
    tw = set(sps)
    say(f"c3: oracle over {len(tw)} coefficient blocks")
    tables = helpers.sparse_tables(ctx, [tw, tw, set(), set()])
    orc = O.Oracle(op, tables)
    want = np.array([orc.green_pair(0, 1, *qi[k, 1:7], qt[k, 0], qt[k, 1]) for k in samp])
    assert bits_equal(ex[samp], want)
    # FP64 matrix-core form: the contraction sums nu*nv products in another
    # order; both forms are within 2 (nu + nv) u sum|c_ab| of the exact sum
    # (|T_k| <= 1 on [-1, 1]), which bounds their difference
    blk = w["nc"] * w["nc"]
    u = 2.0 ** -53
    worst = 0.0
    for j, k in enumerate(samp):
        d = O.descriptor(int(qi[k, 1]), int(qi[k, 2]), qt[k, 0], qt[k, 1], *[int(x) for x in qi[k, 3:7]],
                         w["beta"], 1e-13, w["nc"], w["nc"], N1, N2, qt[k, 0] > qt[k, 1])
        if d["corner"]:
            assert bits_equal(mf[k], ex[k])
            continue
        br = tables[0].upper if d["upper"] else tables[0].lower
        abs_sum = float(np.sum(np.abs(br[d["base"]:d["base"] + blk])))
        bound = 2 * 2 * (2 * w["nc"]) * u * abs_sum
        worst = max(worst, float(np.max(np.abs(mf[k] - ex[k]))) / bound)
    say(f"c3: max |mfma64 - exact| / summation bound = {worst:.3e}")
    assert worst <= 1.0
    rel = np.abs(mf - ex) / np.maximum(np.abs(ex), 1e-12 * np.max(np.abs(ex)))
    say(f"c3: max rel err (floor 1e-12 max|G|) over 2^20 = {np.max(rel):.3e}, mean {np.mean(rel):.3e}")
    assert np.max(np.abs(mf - ex)) < 1e-12
    # the table against the direct Green function at 48 sampled lookups
    calc = O.Calc(N1, N2, w["beta"], w["mu"] - w["alpha"], True, w["mask"], w["radius"])
    err = 0.0
    for k in samp[:48]:
        t1, t2 = qt[k]
        if t1 == t2:
            continue
        g = helpers.direct_twisted_entry(calc, *[int(x) for x in qi[k, 1:7]], t1, t2, t1 > t2, N1)
        err = max(err, abs(ex[k, 0] - g))
    say(f"c3: max |table - direct| at 48 lookups = {err:.3e}")
    assert err < 1e-9
    ctx.close()


# ---------------------------------------------------------------------------
# C4
# ---------------------------------------------------------------------------
def test_c4_envelopes_bit_identical(gpu_available):
    w = bench.C4
    op = oracle_params(w)
    ctx = gpu_context(w, op)
    say("c4: building tables (6x6, 128 coefficients)")
    ctx.build_tables(pbc=True, mask=w["mask"], mask_param=w["width"])
    orders, verts = rng.synthetic_configs(w["base_seed"], 0xC4, w["chains"], 2, w["envelope"], squared=True)
    say(f"c4: {len(verts)} envelopes, orders {dict(zip(*np.unique(orders, return_counts=True)))}")
    out, ob, st = run_plan(ctx, verts, "exact")
    assert not np.any(st & 1)
    idx = pick_per_order(orders, 2, extra={11: 4, 12: 4})
    assert set(int(orders[i]) for i in idx) == set(int(x) for x in orders)
    configs = [verts[i] for i in idx]
    orc = sparse_oracle(ctx, op, configs)
    for i in idx:
        say(f"c4: envelope {i} (order {orders[i]}) against the oracle")
        assert bits_equal(out[ob[i]:ob[i + 1]], orc.all_subsets(verts[i])), (i, orders[i])
    ctx.close()


# ---------------------------------------------------------------------------
# C5
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c5():
    w = bench.C5
    op = oracle_params(w)
    ctx = gpu_context(w, op)
    say("c5: building tables on the GPU (10x10, 192 coefficients, node chunks 7 Nc + 3)")
    ctx.build_tables(pbc=True, mask=w["mask"], mask_param=w["width"], node_chunk=7 * w["nc"] + 3)
    orders, verts = rng.synthetic_configs(w["base_seed"], w["tag"], w["configs"], w["order_min"],
                                          w["order_max"], squared=True)
    yield dict(w=w, ctx=ctx, op=op, orders=orders, verts=verts)
    ctx.close()


def test_c5_high_orders_bit_identical(gpu_available, c5):
    ctx, orders, verts = c5["ctx"], c5["orders"], c5["verts"]
    assert int(orders.max()) == 14
    say("c5: 1024 configurations, exact")
    out, ob, st = run_plan(ctx, verts, "exact")
    assert not np.any(st & 1)
    idx = pick_per_order(orders, 1, extra={13: 2, 14: 2})
    orc = sparse_oracle(ctx, c5["op"], [verts[i] for i in idx])
    for i in idx:
        say(f"c5: configuration {i} (order {orders[i]}) against the oracle")
        assert bits_equal(out[ob[i]:ob[i + 1]], orc.all_subsets(verts[i])), (i, orders[i])


def test_c5_all_subset_full_equals_single_order(gpu_available, c5):
    # tests/test_integrand_all_subsets.cpp:41-62 at orders 13 and 14: the
    # all-subset value of the full mask and the single-order integrand agree
    # (2e-9, absolute below 1, relative above)
    ctx, orders, verts = c5["ctx"], c5["orders"], c5["verts"]
    idx = [int(i) for n in (13, 14) for i in np.nonzero(orders == n)[0][:3]]
    configs = [verts[i] for i in idx]
    say(f"c5: all-subset[full] vs single-order, {len(idx)} configurations at orders 13-14")
    allv, ob, _ = run_plan(ctx, configs, "exact")
    single, _, _ = run_plan(ctx, configs, "exact", single=True)
    for k, v in enumerate(configs):
        full = allv[ob[k + 1] - 1]
        d = single[k]
        assert abs(full - d) <= 2e-9 * max(1.0, abs(full), abs(d)), (idx[k], full, d)


def test_c5_tables_against_direct_green_function(gpu_available, c5):
    # the 192-coefficient tables (made by the FP64 producer and the
    # matrix-core cosine transform in uneven node chunks) against the direct
    # graded two-replica Green function at seeded times and sites
    import torch
    w = c5["w"]
    ctx = c5["ctx"]
    N1, N2 = w["N1"], w["N2"]
    r = np.random.default_rng(55)
    n = 24
    qi = np.zeros((n, 7), dtype=np.int32)
    qi[:, 1:3] = r.integers(0, 2, (n, 2))
    qi[:, 3] = r.integers(0, N1, n)
    qi[:, 4] = r.integers(0, N2, n)
    qi[:, 5] = r.integers(0, N1, n)
    qi[:, 6] = r.integers(0, N2, n)
    qt = r.uniform(0.0, w["beta"], (n, 2))
    d_qi = torch.tensor(qi, device="cuda")
    d_qt = torch.tensor(qt, device="cuda")
    ex = green_pair_eval(ctx, d_qi, d_qt, n, "exact")
    calc = O.Calc(N1, N2, w["beta"], w["mu"] - w["alpha"], True, w["mask"], w["width"])
    err = 0.0
    for k in range(n):
        t1, t2 = qt[k]
        g = helpers.direct_twisted_entry(calc, *[int(x) for x in qi[k, 1:7]], t1, t2, t1 > t2, N1)
        err = max(err, abs(ex[k, 0] - g))
    say(f"c5: max |table - direct| at {n} lookups = {err:.3e}")
    assert err < 1e-9


# ---------------------------------------------------------------------------
# FP64 matrix-core form against the first-order error bound (evaluation kind
# 3): at the headline workload (C2 order 10) and the largest (C5, orders up to
# 14) every value of the matrix-core form is within twice the bound of the
# bit-identical FP64 form, and so is the FP64 form on tables moved by one ulp.
# ---------------------------------------------------------------------------
def bound_report(got, ref, bound):
    """(max |got - ref| / bound, max relative error over the values the bound
    resolves to 1e-9 (|ref| >= 2e9 bound), fraction of nonzero values so
    resolved).  At half filling the odd-order connected values vanish
    identically: they sit at the round-off floor, far below their bound, and
    no relative error is defined for them."""
    ratio = np.abs(got - ref) / np.maximum(bound, 1e-300)
    nz = ref != 0
    res = nz & (np.abs(ref) >= 2e9 * bound)
    rel = np.abs(got - ref)[res] / np.abs(ref)[res]
    return (float(np.max(ratio)), float(np.max(rel)) if rel.size else 0.0,
            float(np.count_nonzero(res) / max(1, np.count_nonzero(nz))))


def test_c2_mfma64_within_error_bound(gpu_available, c2):
    orders, verts = c2["orders"], c2["verts"]
    idx = list(np.nonzero(orders == 10)[0])
    configs = [verts[i] for i in idx]
    say(f"c2: {len(configs)} order-10 configurations, exact / mfma64 / bound")
    ex, _, _ = run_plan(c2["ctx"], configs, "exact")
    mf, _, _ = run_plan(c2["ctx"], configs, "mfma64")
    bd, _, _ = run_plan(c2["ctx"], configs, "exact", kind=3)
    assert np.all(np.isfinite(bd))
    r_mf, rel_mf, frac_mf = bound_report(mf, ex, bd)
    # one-ulp control: the bit-identical form on tables with every coefficient
    # moved by one ulp in a seeded random direction
    r = np.random.default_rng(7)
    made = {}
    pt = []
    for t in c2["tables"]:
        if id(t) not in made:
            made[id(t)] = O.Table(*[np.nextafter(a, np.where(r.random(a.shape) < 0.5, -np.inf, np.inf))
                                    for a in (t.upper, t.lower)], t.ucorner, t.lcorner)
        pt.append(made[id(t)])
    w = c2["w"]
    ctx_ulp = gpu_context(w, c2["op"])
    ctx_ulp.upload_tables(pt)
    ul, _, _ = run_plan(ctx_ulp, configs, "exact")
    ctx_ulp.close()
    r_ul, rel_ul, frac_ul = bound_report(ul, ex, bd)
    say(f"c2 order 10: mfma64 max err/bound {r_mf:.3e}, max rel on resolved values {rel_mf:.3e} "
        f"(resolved fraction {frac_mf:.4f}); one-ulp control max err/bound {r_ul:.3e}, max rel {rel_ul:.3e}")
    assert r_mf <= 2.0
    assert r_ul <= 2.0
    assert rel_mf <= 1e-9 and frac_mf > 0.0


def test_c2_tree_order_sums(gpu_available, c2):
    # The non-exact modes sum each connected value's terms in tree order
    # (cdet_ctx_set_sum_order, default CDET_SUM_TREE).  At order 10 popcount
    # layers 8, 9 and 10 run 8, 8 and 64 lanes per mask and the layers below
    # keep the serial chain, so masks of at most seven vertices equal the
    # reference-order values bit for bit and the others move at round-off.
    # Both orders stay within twice the first-order bound of the FP64 values;
    # the exact mode ignores the setting.
    ctx = c2["ctx"]
    orders, verts = c2["orders"], c2["verts"]
    idx = list(np.nonzero(orders == 10)[0])[:256]
    configs = [verts[i] for i in idx]
    say(f"c2: {len(configs)} order-10 configurations, tree-order and reference-order sums")
    ex, _, _ = run_plan(ctx, configs, "exact")
    bd, _, _ = run_plan(ctx, configs, "exact", kind=3)
    tree, _, _ = run_plan(ctx, configs, "mfma64")
    ctx.set_sum_order("reference")
    try:
        serial, _, _ = run_plan(ctx, configs, "mfma64")
        ex_ref, _, _ = run_plan(ctx, configs, "exact")
    finally:
        ctx.set_sum_order("tree")
    assert bits_equal(ex, ex_ref)
    pop = np.array([bin(m).count("1") for m in range(1 << 10)])
    t2, s2 = tree.reshape(len(configs), -1), serial.reshape(len(configs), -1)
    assert bits_equal(t2[:, pop <= 7], s2[:, pop <= 7])
    moved = np.count_nonzero(t2[:, pop >= 8] != s2[:, pop >= 8])
    r_tree, rel_tree, _ = bound_report(tree, ex, bd)
    r_ser, rel_ser, _ = bound_report(serial, ex, bd)
    say(f"c2 order 10: {moved} of {t2[:, pop >= 8].size} values of popcount >= 8 moved; max err/bound "
        f"tree {r_tree:.3e} serial {r_ser:.3e}; max rel on resolved values tree {rel_tree:.3e} serial {rel_ser:.3e}")
    assert moved > 0
    assert r_tree <= 2.0 and r_ser <= 2.0
    assert rel_tree <= 1e-9 and rel_ser <= 1e-9


def test_c2_single_order_mfma64_within_error_bound(gpu_available, c2):
    # Single-order integrands (kind 1) in the FP64 matrix-core mode, whose
    # single-anchor recursion sums in tree order: the connected value of the
    # full mask is formed by the same chains as in the all-subset recursion
    # (masks holding bit 0 have it as their anchor in both), so each value is
    # within twice the first-order bound of the all-subset value of the full
    # mask (kind 3) of the bit-identical FP64 value; the prefactors differ by
    # their last-bit rounding only.  Orders up to 4 have no tree-order layer:
    # there the reference order gives the same bits.
    ctx = c2["ctx"]
    orders, verts = c2["orders"], c2["verts"]
    idx = [i for i in pick_per_order(orders, 6, extra={10: 24}) if orders[i] >= 2]
    configs = [verts[i] for i in idx]
    say(f"c2: {len(configs)} single-order configurations (orders 2..10), exact / mfma64 / bound")
    ex, _, _ = run_plan(ctx, configs, "exact", single=True)
    mf, _, _ = run_plan(ctx, configs, "mfma64", single=True)
    bd, ob, _ = run_plan(ctx, configs, "exact", kind=3)
    ctx.set_sum_order("reference")
    try:
        serial, _, _ = run_plan(ctx, configs, "mfma64", single=True)
    finally:
        ctx.set_sum_order("tree")
    full = np.array([bd[ob[k] + (1 << len(v)) - 1] for k, v in enumerate(configs)])
    ratio = np.abs(mf - ex) / np.maximum(2.0 * full + 1e-14 * np.abs(ex), 1e-300)
    n = np.array([len(v) for v in configs])
    say(f"c2 single-order: max |mfma64 - exact| / (2 bound) {ratio.max():.3e}; "
        f"{int(np.count_nonzero(mf[n >= 8] != serial[n >= 8]))} of {int(np.count_nonzero(n >= 8))} "
        "orders >= 8 moved by the tree-order sums")
    assert np.all(np.isfinite(mf))
    assert ratio.max() <= 1.0
    assert bits_equal(mf[n <= 4], serial[n <= 4])


def test_c5_mfma64_within_error_bound(gpu_available, c5):
    ctx, orders, verts = c5["ctx"], c5["orders"], c5["verts"]
    idx = pick_per_order(orders, 1, extra={12: 2, 13: 3, 14: 3})
    configs = [verts[i] for i in idx]
    say(f"c5: {len(configs)} configurations (orders {sorted(set(int(orders[i]) for i in idx))}), "
        "exact / mfma64 / bound")
    ex, ob, _ = run_plan(ctx, configs, "exact")
    mf, _, _ = run_plan(ctx, configs, "mfma64")
    bd, _, _ = run_plan(ctx, configs, "exact", kind=3)
    assert np.all(np.isfinite(bd))
    for k, i in enumerate(idx):
        s = slice(ob[k], ob[k + 1])
        r_mf, rel_mf, frac = bound_report(mf[s], ex[s], bd[s])
        say(f"c5 order {orders[i]}: mfma64 max err/bound {r_mf:.3e}, max rel on resolved values {rel_mf:.3e} "
            f"(resolved fraction {frac:.4f})")
        assert r_mf <= 2.0, (i, orders[i])
        assert rel_mf <= 1e-9, (i, orders[i])
    # single-order integrands of the same configurations in the matrix-core
    # mode (order 14: the connected values in LDS and the products in global
    # scratch, tree-order sums) against the bound of the full mask, as at C2
    ex1, _, _ = run_plan(ctx, configs, "exact", single=True)
    mf1, _, _ = run_plan(ctx, configs, "mfma64", single=True)
    full = np.array([bd[ob[k + 1] - 1] for k in range(len(configs))])
    ratio = np.abs(mf1 - ex1) / np.maximum(2.0 * full + 1e-14 * np.abs(ex1), 1e-300)
    say(f"c5 single-order: max |mfma64 - exact| / (2 bound) {ratio.max():.3e}")
    assert np.all(np.isfinite(mf1))
    assert ratio.max() <= 1.0

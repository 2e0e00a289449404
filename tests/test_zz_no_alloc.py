"""The no-allocation entry point (cdet_eval_all_subsets_noalloc) against the
allocating one (cdet_eval_all_subsets): same values bit for bit, on calls in
a row of decreasing size that share the library's static device scratch, and
on replays of a captured HIP graph."""
import ctypes
import sys

import numpy as np
import pytest

import cdet_amd
import helpers

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits_equal(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def make_ctx(name="c1"):
    s = helpers.SETS[name]
    op = helpers.params_of(name)
    ctx = cdet_amd.Context(cdet_amd.make_params(s["N1"], s["N2"], s["nc"], s["beta"], s["U"], s["mu"],
                                                s["alpha_up"], s["alpha_down"],
                                                derived=(op.scale, op.diag_up_shift, op.diag_down_shift)))
    ctx.upload_tables(helpers.tables_of(name))
    return ctx


def noalloc(ctx, d_v, n, d_out, d_status, stream):
    return cdet_amd.lib().cdet_eval_all_subsets_noalloc(
        ctx.handle, ctypes.c_void_p(d_v.data_ptr()), n, ctypes.c_void_p(d_out.data_ptr()),
        ctypes.c_void_p(d_status.data_ptr()) if d_status is not None else None, ctypes.c_void_p(stream))


def test_noalloc_matches_allocating_entry_point(gpu_available):
    import torch
    ctx = make_ctx("c1")
    stream = torch.cuda.current_stream().cuda_stream
    # largest order first (global-scratch forms), then smaller inputs in the
    # same static buffer, which keeps the previous call's contents
    for k, n in enumerate((14, 10, 6, 2)):
        v = helpers.seeded_configs(12000 + n, [n])[0]
        want = ctx.evaluate_all_subset_integrands(v)
        d_v = torch.tensor(v, dtype=torch.float64, device="cuda")
        d_out = torch.full((1 << n,), float("nan"), dtype=torch.float64, device="cuda")
        d_st = torch.full((1,), 7, dtype=torch.int32, device="cuda")
        say(f"no-allocation call {k}: order {n}")
        rc = noalloc(ctx, d_v, n, d_out, d_st, stream)
        assert rc == 0, cdet_amd._err()
        torch.cuda.synchronize()
        assert bits_equal(d_out.cpu().numpy(), want), n
        assert int(d_st.cpu()[0]) & 1 == 0
    # fallback path (TOL_FPM 0.9: most pivots fall back) through the static
    # buffer, status bit 1 reported
    ctx.close()
    s = helpers.SETS["c1"]
    op = helpers.params_of("c1")
    fb = cdet_amd.Context(cdet_amd.make_params(s["N1"], s["N2"], s["nc"], s["beta"], s["U"], s["mu"],
                                               s["alpha_up"], s["alpha_down"], tol_fpm=0.9,
                                               derived=(op.scale, op.diag_up_shift, op.diag_down_shift)))
    fb.upload_tables(helpers.tables_of("c1"))
    v = helpers.seeded_configs(12100, [7])[0]
    want = fb.evaluate_all_subset_integrands(v)
    d_v = torch.tensor(v, dtype=torch.float64, device="cuda")
    d_out = torch.empty(1 << 7, dtype=torch.float64, device="cuda")
    d_st = torch.zeros(1, dtype=torch.int32, device="cuda")
    say("no-allocation call with the FPM fallback")
    assert noalloc(fb, d_v, 7, d_out, d_st, stream) == 0, cdet_amd._err()
    torch.cuda.synchronize()
    assert bits_equal(d_out.cpu().numpy(), want)
    assert int(d_st.cpu()[0]) & 2
    fb.close()


def test_noalloc_rejects_sizes_and_orders(gpu_available):
    import torch
    ctx = make_ctx("ref2x2")
    d_v = torch.zeros((16, 4), dtype=torch.float64, device="cuda")
    d_out = torch.zeros(1 << 15, dtype=torch.float64, device="cuda")
    rc = noalloc(ctx, d_v, 15, d_out, None, 0)
    assert rc == -6  # CDET_ERR_CAPACITY
    assert "CDET_MAX_ORDER" in cdet_amd._err()
    rc = noalloc(ctx, d_v, 1, d_out, None, 0)
    assert rc == -2  # CDET_ERR_ORDER
    assert "envelope order >= 2" in cdet_amd._err()
    ctx.close()


def test_noalloc_in_captured_graph(gpu_available):
    # Nothing in the call allocates or clears with memset, so it can be
    # recorded into a HIP graph; replays, with eager work and synchronisation
    # between them, give the eager values.
    import torch
    ctx = make_ctx("c1")
    n = 9
    v = helpers.seeded_configs(12200, [n])[0]
    want = ctx.evaluate_all_subset_integrands(v)
    d_v = torch.tensor(v, dtype=torch.float64, device="cuda")
    d_out = torch.full((1 << n,), float("nan"), dtype=torch.float64, device="cuda")
    d_st = torch.zeros(1, dtype=torch.int32, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        assert noalloc(ctx, d_v, n, d_out, d_st, side.cuda_stream) == 0, cdet_amd._err()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    assert bits_equal(d_out.cpu().numpy(), want)
    g = torch.cuda.CUDAGraph()
    say("capturing the no-allocation call")
    with torch.cuda.graph(g):
        rc = noalloc(ctx, d_v, n, d_out, d_st, torch.cuda.current_stream().cuda_stream)
    assert rc == 0, cdet_amd._err()
    for r in range(3):
        d_out.fill_(float("nan"))
        other = ctx.evaluate_all_subset_integrands(helpers.seeded_configs(12300 + r, [6])[0])
        assert np.all(np.isfinite(other))
        torch.cuda.synchronize()
        say(f"graph replay {r}")
        g.replay()
        torch.cuda.synchronize()
        assert bits_equal(d_out.cpu().numpy(), want), r
    ctx.close()

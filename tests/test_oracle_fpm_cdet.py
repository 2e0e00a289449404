"""Oracle stage checks for the fast principal minors and the CDet recursions.

Restates the checks of the reference's tests/test_fpm_cdet.cpp (same seed,
same input recipe, same tolerances) and adds independent ones: numpy
determinants of every principal submatrix and an independent set-partition
cumulant.
"""
import itertools
import math

import numpy as np
import pytest

from oracle import oracle as O


def close(a, e, abs_tol, rel_tol):
    # tests/audit_test_utils.h:18-30: scale = max(1, |a|, |e|)
    scale = max(1.0, abs(a), abs(e))
    return math.isfinite(a) and math.isfinite(e) and abs(a - e) <= abs_tol + rel_tol * scale


def reference_stream():
    """mt19937_64(0x8ac341ee) with uniform_real_distribution(-0.8, 0.8),
    consumed in the order of tests/test_fpm_cdet.cpp:141-183."""
    return iter(O.uniform_stream(0x8AC341EE, -0.8, 0.8, 200000))


def np_minors(A):
    n = A.shape[0]
    out = np.zeros((1 << n) - 1)
    for m in range(1, 1 << n):
        idx = [i for i in range(n) if m >> i & 1]
        out[m - 1] = np.linalg.det(A[np.ix_(idx, idx)])
    return out


def test_fpm_matches_pivoted_determinants_reference_matrices():
    it = reference_stream()
    for n in range(1, 9):
        A = np.array([next(it) for _ in range(n * n)]).reshape(n, n)
        for d in range(n):
            A[d, d] += 1.2 + 0.1 * d
        brute = O.det_minors(A)
        indep = np_minors(A)
        assert np.allclose(brute, indep, rtol=1e-12, atol=1e-12)
        for tol, shift in ((0.0, -1.0), (0.0, 0.37), (1.0e-5, -1.0)):
            pm, fb = O.mat2pm(A, tol, shift)
            assert fb == 0
            for x, y in zip(pm, brute):
                assert close(x, y, 2e-10, 2e-10)


def test_fpm_pivot_pathology_takes_fallback():
    A = np.array([[0.0, 1.0, 0.2], [1.0, 0.3, -0.1], [0.2, -0.1, 1.4]])
    pm, fb = O.mat2pm(A, 1.0e-5, 0.0)
    assert fb == 1
    brute = O.det_minors(A)
    for x, y in zip(pm, brute):
        assert close(x, y, 2e-13, 2e-13)
    assert np.allclose(pm, np_minors(A), atol=1e-14)


def test_fpm_bit_mask_convention():
    # pm[mask-1] = det(A[mask, mask]), bit i <-> row i: diagonal matrix
    n = 5
    d = np.array([2.0, 3.0, 5.0, 7.0, 11.0])
    pm, _ = O.mat2pm(np.diag(d))
    for m in range(1, 1 << n):
        want = np.prod([d[i] for i in range(n) if m >> i & 1])
        assert abs(pm[m - 1] - want) <= 1e-12 * want


def partition_cumulant(mask, moments):
    """Connected part by inclusion over set partitions:
    kappa(S) = sum_pi (-1)^(|pi|-1) (|pi|-1)! prod_{B in pi} m(B)."""
    elems = [i for i in range(64) if mask >> i & 1]

    def partitions(items):
        if not items:
            yield []
            return
        first, rest = items[0], items[1:]
        for k in range(len(rest) + 1):
            for combo in itertools.combinations(rest, k):
                block = (first,) + combo
                remaining = [x for x in rest if x not in combo]
                for p in partitions(remaining):
                    yield [block] + p

    total = 0.0
    for p in partitions(elems):
        k = len(p)
        prod = 1.0
        for b in p:
            prod *= moments[sum(1 << i for i in b)]
        total += (-1) ** (k - 1) * math.factorial(k - 1) * prod
    return total


def test_ztransform_matches_partition_cumulant():
    it = reference_stream()
    for n in range(1, 9):  # advance past the FPM matrices (tests/test_fpm_cdet.cpp:144-156)
        for _ in range(n * n):
            next(it)
    for order in range(1, 8):
        sc = 1 << order
        up = np.zeros(sc - 1)
        dn = np.zeros(sc - 1)
        moments = np.zeros(sc)
        for m in range(1, sc):
            up[m - 1] = 0.6 + 0.08 * next(it)
            dn[m - 1] = 0.7 + 0.08 * next(it)
            moments[m] = up[m - 1] * dn[m - 1]
        dp = O.ztransform_all(up, dn, order, 0.0, 0.0)
        for m in range(1, sc):
            assert close(dp[m], partition_cumulant(m, moments), 2e-11, 2e-11)
        # single-anchor form agrees with the all-subset value of the full set
        single = O.ztransform_single(up, dn, order, 0.0, 0.0)
        assert close(single, dp[sc - 1], 1e-13, 1e-13)


def test_roundoff_filter_cancels_constant_moment():
    for order in range(2, 10):
        sc = 1 << order
        ones = np.ones(sc - 1)
        dp = O.ztransform_all(ones, ones, order, 2.0e-16, 1.0e4)
        assert abs(dp[sc - 1]) < 1e-14
        # without the filter the cancellation leaves round-off (or exact 0)
        raw = O.ztransform_all(ones, ones, order, 0.0, 0.0)
        assert abs(raw[sc - 1]) < 1e-9


def test_branch_epsilon_corner_selection():
    # tests/test_fpm_cdet.cpp:116-128
    uc = O.descriptor(0, 1, 0.5e-13, 0.25e-13, 0, 0, 0, 0, 8.0, 1.0e-13, 25, 25, 3, 3, True)
    ur = O.descriptor(0, 1, 2.0e-13, 1.0e-13, 0, 0, 0, 0, 8.0, 1.0e-13, 25, 25, 3, 3, True)
    lc = O.descriptor(0, 1, 0.25e-13, 0.5e-13, 0, 0, 0, 0, 8.0, 1.0e-13, 25, 25, 3, 3, False)
    assert uc["corner"] and lc["corner"] and not ur["corner"]


def test_descriptor_index_and_mapping():
    # sitepair = ((((ra*2+rb)*N1+x1)*N2+y1)*N1+x2)*N2+y2, base = sitepair*nu*nv
    d = O.descriptor(1, 0, 3.0, 1.0, 2, 1, 0, 2, 8.0, 1e-13, 5, 7, 3, 4, True)
    sp = ((((1 * 2 + 0) * 3 + 2) * 4 + 1) * 3 + 0) * 4 + 2
    assert d["base"] == sp * 35
    assert d["u"] == 2.0 * 3.0 / 8.0 - 1.0 and d["v"] == 2.0 * 1.0 / 3.0 - 1.0
    lo = O.descriptor(0, 0, 1.0, 3.0, 0, 0, 0, 0, 8.0, 1e-13, 5, 5, 3, 3, False)
    assert lo["u"] == 2.0 * 3.0 / 8.0 - 1.0 and lo["v"] == 2.0 * 1.0 / 3.0 - 1.0
    # clamping at the triangle edge
    e = O.descriptor(0, 0, 8.0, 8.0, 0, 0, 0, 0, 8.0, 1e-13, 5, 5, 3, 3, True)
    assert e["u"] == 1.0 and e["v"] == 1.0

"""Oracle checks of the assembled integrand."""
import math
from fractions import Fraction

import numpy as np

import helpers
from oracle import oracle as O


def test_all_subset_equals_single_order_reference_envelopes():
    # tests/test_integrand_all_subsets.cpp:41-62: five 6-vertex envelopes,
    # mt19937_64(0x4dcd173b), uniform(0.01, 0.99), tolerance 2e-9.
    orc = helpers.oracle_of("ref2x2")
    u = O.uniform_stream(0x4DCD173B, 0.01, 0.99, 5 * 6 * 4).reshape(5, 6, 4)
    for env in u:
        allv = orc.all_subsets(env)
        assert allv.shape == (64,) and allv[0] == 0.0
        for m in range(1, 64):
            sel = env[[i for i in range(6) if m >> i & 1]]
            d = orc.single(sel)
            scale = max(1.0, abs(allv[m]), abs(d))
            assert abs(allv[m] - d) <= 2e-9 + 2e-9 * scale


def test_order_errors_match_reference_messages():
    orc = helpers.oracle_of("ref2x2")
    v = np.full((1, 4), 0.5)
    try:
        orc.all_subsets(v)
        raise AssertionError("expected failure")
    except RuntimeError as e:
        assert "envelope order >= 2" in str(e)


# ---------------------------------------------------------------------------
# Exact pin of the assembly: tables holding one constant per class make every
# vertex matrix s (c J + d I) (J all ones, s = sqrt(U N1 N2 beta), d the spin's
# diagonal shift), whose k x k principal minors are s^k d^(k-1) (d + k c).  The
# all-subset recursion then depends on k only, and every value follows in exact
# rational arithmetic: the prefactor (-rc)^k / (k! N1 N2)
# (include/cdet/integrand_replica_randomly_sampled.h:155-171), the singleton
# -rc U beta (d_up d_down - alpha_up alpha_down / U^2) (:833-877) and the
# combination twisted - 2 ordinary (include/cdet/cdet_integrand_evaluator.h:262,
# :349).  A wrong factorial, sign, lattice normalisation, singleton term or
# combination weight moves a value by a factor, far outside the tolerance.
# ---------------------------------------------------------------------------
EXACT = dict(N1=2, N2=2, beta=1.0, U=1.0, mu=0.0, alpha_up=0.5, alpha_down=0.25, nc=4)


def constant_tables(c_tw, c_ord, N1=2, N2=2, nc=4):
    SP = 4 * (N1 * N2) ** 2

    def table(c):
        co = np.zeros((SP, nc, nc))
        co[:, 0, 0] = c  # T_0(u) T_0(v) = 1: the table is c everywhere
        return O.Table(co.ravel(), co.ravel(), np.full(SP, c), np.full(SP, c))

    tw, od = table(c_tw), table(c_ord)
    return [tw, tw, od, od]


def exact_values(n, c_tw, c_ord, w=EXACT):
    """Exact all-subset values by subset size k (values depend on k only)."""
    U, beta = Fraction(w["U"]), Fraction(w["beta"])
    N1, N2 = w["N1"], w["N2"]
    s2 = U * N1 * N2 * beta  # s^2, a perfect square here
    s = Fraction(math.isqrt(s2.numerator), math.isqrt(s2.denominator))
    assert s * s == s2
    a_up, a_dn = Fraction(w["alpha_up"]), Fraction(w["alpha_down"])
    d_up, d_dn = a_dn / U, a_up / U  # diagonal shifts of the up / down matrices

    def cls(c, rc):
        c = Fraction(c)
        pm = lambda d, k: s ** k * d ** (k - 1) * (d + k * c)  # noqa: E731
        a = {k: pm(d_up, k) * pm(d_dn, k) for k in range(1, n + 1)}
        dp, mag = {1: a[1]}, {1: abs(a[1])}
        for k in range(2, n + 1):
            terms = [math.comb(k - 1, j - 1) * dp[j] * a[k - j] for j in range(1, k)]
            dp[k] = a[k] - sum(terms)
            mag[k] = abs(a[k]) + sum(abs(t) for t in terms)
        val = {1: -rc * U * beta * ((c + d_up) * (c + d_dn) - a_up * a_dn / (U * U))}
        scl = {1: abs(val[1])}
        for k in range(2, n + 1):
            pref = Fraction((-rc) ** k, math.factorial(k) * N1 * N2)
            val[k] = pref * dp[k]
            scl[k] = abs(pref) * mag[k]
        return val, scl

    vt, st = cls(c_tw, 2)
    vo, so = cls(c_ord, 1)
    return ({k: vt[k] - 2 * vo[k] for k in vt}, {k: st[k] + 2 * so[k] for k in st})


def exact_oracle(c_tw, c_ord):
    w = EXACT
    p = O.make_params(w["N1"], w["N2"], w["nc"], w["beta"], w["U"], w["mu"], w["alpha_up"], w["alpha_down"])
    assert p.scale == 2.0 and p.diag_up_shift == 0.25 and p.diag_down_shift == 0.5
    return O.Oracle(p, constant_tables(c_tw, c_ord))


def test_prefactor_rationals():
    # scale and diagonal shifts are correctly rounded
    p = O.make_params(4, 4, 8, 6.0, 2.0, 1.0, 1.0, 1.0)
    assert p.scale == math.sqrt(192.0)
    p = O.make_params(2, 2, 8, 2.5, 1.7, 0.3, 0.55, 0.4)
    assert p.diag_up_shift == 0.4 / 1.7 and p.diag_down_shift == 0.55 / 1.7
    assert p.scale == math.sqrt(Fraction(1.7) * 4 * Fraction(2.5))


def test_assembly_exact_rationals_all_subsets():
    c_tw, c_ord = 0.25, -0.125
    orc = exact_oracle(c_tw, c_ord)
    n = 9
    want, scale = exact_values(n, c_tw, c_ord)
    v = helpers.seeded_configs(4242, [n])[0]
    got = orc.all_subsets(v)
    assert got[0] == 0.0
    for m in range(1, 1 << n):
        k = bin(m).count("1")
        assert abs(Fraction(got[m]) - want[k]) <= Fraction(1e-13) * scale[k], (m, k, got[m], float(want[k]))
    # the pin is not vacuous: the connected values are far from zero
    assert all(abs(want[k]) > Fraction(1e-3) * scale[k] for k in (2, 4, 6, 8))


def test_assembly_exact_rationals_single_order():
    c_tw, c_ord = 0.25, -0.125
    orc = exact_oracle(c_tw, c_ord)
    want, scale = exact_values(8, c_tw, c_ord)
    for n in range(1, 9):
        v = helpers.seeded_configs(4343 + n, [n])[0]
        got = orc.single(v)
        assert abs(Fraction(got) - want[n]) <= Fraction(1e-13) * scale[n], (n, got, float(want[n]))


def test_assembly_exact_rationals_detect_planted_prefactor_error():
    # a factorial off by one (k! -> (k-1)!) or a dropped N1 N2 moves every
    # value of order >= 2 by a factor >= 2: the bound above rejects it
    want, scale = exact_values(8, 0.25, -0.125)
    for k in range(2, 9):
        for wrong in (want[k] * k, want[k] * 4):
            assert abs(wrong - want[k]) > Fraction(1e-13) * scale[k]


def test_spin_symmetric_tables_give_identical_spin_matrices():
    # ALPHA_UP == ALPHA_DOWN: up and down vertex matrices are bit-identical
    # (include/cdet/load_data.h:70-91 with the alpha_down/U, alpha_up/U shifts).
    orc = helpers.oracle_of("ref2x2_eq")
    v = helpers.seeded_configs(11, [5])[0]
    for which in (0, 1):
        mu, md, du, dd = orc.vertex_matrices(v, which)
        assert np.array_equal(mu, md) and np.array_equal(du, dd)

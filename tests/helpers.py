"""Shared fixtures: seeded inputs and oracle-built tables (tests only)."""
import functools

import numpy as np

from oracle import oracle as O

# Parameter sets.  "ref2x2" is the reference's all-subset integrand test
# (tests/test_integrand_all_subsets.cpp:11-39); "c1" the 3x3 workload
# (t=1, half filling, beta=4, one-column subsystem, 64 coefficients);
# "c1_neq" the same with unequal spin shifts (distinct up and down tables);
# "p8x8" the largest lattice of the LDS producer (64 sites) at a small
# coefficient count; "c2" the 4x4 workload (beta=6, 2x2 corner block, 96 coefficients).
SETS = {
    "ref2x2": dict(N1=2, N2=2, beta=2.5, U=1.7, mu=0.3, alpha_up=0.55, alpha_down=0.4, nc=18,
                   mask="first_row", mask_param=0),
    "ref2x2_eq": dict(N1=2, N2=2, beta=2.5, U=1.7, mu=0.3, alpha_up=0.55, alpha_down=0.55, nc=18,
                      mask="first_row", mask_param=0),
    "c1": dict(N1=3, N2=3, beta=4.0, U=2.0, mu=1.0, alpha_up=1.0, alpha_down=1.0, nc=64,
               mask="first_column", mask_param=0),
    "c1_neq": dict(N1=3, N2=3, beta=4.0, U=2.0, mu=1.0, alpha_up=1.15, alpha_down=0.85, nc=64,
                   mask="first_column", mask_param=0),
    "p8x8": dict(N1=8, N2=8, beta=8.0, U=2.0, mu=1.0, alpha_up=1.2, alpha_down=0.8, nc=16,
                 mask="origin_square", mask_param=2),
    "c2": dict(N1=4, N2=4, beta=6.0, U=2.0, mu=1.0, alpha_up=1.0, alpha_down=1.0, nc=96,
               mask="origin_square", mask_param=1),
    # lattices above 64 sites (the inversion producer), table-level coverage
    # at a small coefficient count: 72 sites with distinct spin shifts, and the
    # 10x10 C5 lattice with its x < 5 half
    "p9x8": dict(N1=9, N2=8, beta=5.0, U=2.0, mu=0.7, alpha_up=1.1, alpha_down=0.9, nc=4,
                 mask="left_columns", mask_param=4),
    "p10x10": dict(N1=10, N2=10, beta=10.0, U=2.0, mu=1.0, alpha_up=1.0, alpha_down=1.0, nc=4,
                   mask="left_columns", mask_param=5),
}


def params_of(name, **over):
    s = dict(SETS[name])
    s.update(over)
    return O.make_params(s["N1"], s["N2"], s["nc"], s["beta"], s["U"], s["mu"], s["alpha_up"], s["alpha_down"])


@functools.lru_cache(maxsize=None)
def tables_of(name):
    s = SETS[name]
    return O.build_tables(s["N1"], s["N2"], s["beta"], s["mu"], s["alpha_up"], s["alpha_down"], s["nc"],
                          mask=s["mask"], mask_param=s["mask_param"])


def oracle_of(name):
    return O.Oracle(params_of(name), tables_of(name))


def seeded_configs(seed, orders, lo=0.0, hi=1.0):
    """Vertex sets: 4 uniform draws per vertex (tau, x, y, replica) from
    mt19937_64(seed) with libstdc++ uniform_real_distribution(lo, hi)."""
    total = int(np.sum(orders))
    u = O.uniform_stream(seed, lo, hi, 4 * total).reshape(total, 4)
    out, k = [], 0
    for n in orders:
        out.append(u[k:k + n].copy())
        k += n
    return out

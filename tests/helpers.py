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


# ---------------------------------------------------------------------------
# Oracle views of large device tables: only the coefficient blocks a sample of
# lookups reads is copied back; the rest of each full-size array stays an
# untouched zero mapping (np.zeros maps pages lazily), so the oracle indexes
# the reference layout directly while host memory holds only those blocks.
# ---------------------------------------------------------------------------
def unit_to_index(coordinate, extent):
    """sampled_coordinate_to_index (include/cdet/g_funcs.h:315-342)."""
    i = int(np.floor(extent * coordinate))
    return min(max(i, 0), extent - 1)


def sitepair_index(ra, rb, x1, y1, x2, y2, N1, N2):
    """replica_sitepair_index (include/cdet/replica_chebyshev_tables.h:365-382)."""
    return ((((ra * 2 + rb) * N1 + x1) * N2 + y1) * N1 + x2) * N2 + y2


def config_sitepairs(verts, N1, N2):
    """Sitepairs the twisted (replicas decoded, rc = 2) and ordinary (replica
    0) classes of a configuration read, as two sets."""
    dec = [(unit_to_index(v[1], N1), unit_to_index(v[2], N2), unit_to_index(v[3], 2)) for v in verts]
    tw, od = set(), set()
    for (x1, y1, r1) in dec:
        for (x2, y2, r2) in dec:
            tw.add(sitepair_index(r1, r2, x1, y1, x2, y2, N1, N2))
            od.add(sitepair_index(0, 0, x1, y1, x2, y2, N1, N2))
    return tw, od


def sparse_tables(ctx, sitepairs_by_kind):
    """Four oracle tables (kinds 0..3) of the context's geometry holding the
    device values of the given sitepair blocks per kind, zeros elsewhere.
    Kinds whose sitepair set is the same object share one table."""
    p = ctx.params
    S = p.N1 * p.N2
    SP = 4 * S * S
    blk = p.cheb_nu * p.cheb_nv
    made = {}
    out = []
    for kind in range(4):
        sps = sitepairs_by_kind[kind]
        if id(sps) in made:
            out.append(made[id(sps)])
            continue
        up = np.zeros(SP * blk)
        lo = np.zeros(SP * blk)
        uc = np.zeros(SP)
        lc = np.zeros(SP)
        idx = np.array(sorted(sps), dtype=np.int64)
        if len(idx):
            bu, bl, cu, cl = ctx.download_blocks(kind, idx)
            for k, sp in enumerate(idx):
                up[sp * blk:(sp + 1) * blk] = bu[k].ravel()
                lo[sp * blk:(sp + 1) * blk] = bl[k].ravel()
            uc[idx] = cu
            lc[idx] = cl
        t = O.Table(up, lo, uc, lc)
        made[id(sps)] = t
        out.append(t)
    return out


def direct_twisted_entry(calc, ra, rb, x1, y1, x2, y2, tau, taup, upper, N1):
    """Replica entry of the graded two-replica Green function from the
    oracle's sector matrices (include/cdet/replica_chebyshev_tables.h:489-510):
    (g0 + g1)/2 on the replica diagonal, +-i (g0 - g1)/2 off it."""
    g0 = calc.sector_matrix(0, tau, taup, upper)
    g1 = calc.sector_matrix(1, tau, taup, upper)
    s1, s2 = y1 * N1 + x1, y2 * N1 + x2
    a, b = g0[s1, s2], g1[s1, s2]
    if ra == rb:
        return ((a + b) / 2).real
    d = a - b
    return ((1j * d) / 2).real if (ra, rb) == (0, 1) else ((-1j * d) / 2).real

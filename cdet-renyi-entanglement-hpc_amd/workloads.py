"""Seeded synthetic workloads of the CDet integrand path on one MI355X.

C2  4x4 periodic Hubbard lattice, t = 1, half filling (U = 2, mu = alpha_up =
    alpha_down = 1), beta = 6, 2x2 corner-block subsystem (origin_square,
    radius 1), 96 Chebyshev coefficients per axis; 8192 configurations with
    orders 1..10 (configuration i from mt19937_64(derive_mcmcmc_seed(
    2026101401, 0xC2, i))).
C3  8x8, beta = 8, 4x4 corner block, 128 coefficients; 2^20 twisted-table
    lookups.
C4  6x6, beta = 8, x < 3 half strip, 128 coefficients; 256 chains with
    envelope orders 2 + floor(11 u^2).
C5  10x10, beta = 10, x < 5 half, 192 coefficients; 1024 configurations at
    orders 2 + floor(13 u^2).
Shared by bench.py, the autotuner (tune.py) and the tests.
"""
import numpy as np

C2 = dict(N1=4, N2=4, beta=6.0, U=2.0, mu=1.0, alpha=1.0, nc=96, mask="origin_square", radius=1,
                configs=8192, order=10, base_seed=2026101401, tag=0xC2)


C3 = dict(N1=8, N2=8, beta=8.0, U=2.0, mu=1.0, alpha=1.0, nc=128, mask="origin_square", radius=2,
          lookups=1 << 20, base_seed=2026101401, tag=0xC3, block=1024)


def c3_queries():
    """2^20 twisted-table lookups on the 8x8 lattice: block i of 1024 lookups
    draws from mt19937_64(derive_mcmcmc_seed(2026101401, 0xC3, i)), per lookup
    six U[0,1) numbers in the order site_i, site_j, replica_a, replica_b, tau,
    tau' (site = floor(64 u) with x = site mod 8, y = site div 8; replica =
    floor(2 u); tau = beta u)."""
    import rng
    w = C3
    nb = w["lookups"] // w["block"]
    eng = rng.MT19937_64([rng.derive_mcmcmc_seed(w["base_seed"], w["tag"], i) for i in range(nb)])
    u = np.stack([eng.uniform() for _ in range(6 * w["block"])], axis=1).reshape(nb, w["block"], 6)
    u = u.reshape(-1, 6)
    S = w["N1"] * w["N2"]
    si = np.minimum((S * u[:, 0]).astype(np.int32), S - 1)
    sj = np.minimum((S * u[:, 1]).astype(np.int32), S - 1)
    qi = np.zeros((len(u), 7), dtype=np.int32)
    qi[:, 1] = np.minimum((2 * u[:, 2]).astype(np.int32), 1)
    qi[:, 2] = np.minimum((2 * u[:, 3]).astype(np.int32), 1)
    qi[:, 3], qi[:, 4] = si % w["N1"], si // w["N1"]
    qi[:, 5], qi[:, 6] = sj % w["N1"], sj // w["N1"]
    qt = np.stack([w["beta"] * u[:, 4], w["beta"] * u[:, 5]], axis=1)
    return qi, qt



C4 = dict(N1=6, N2=6, beta=8.0, U=2.0, mu=1.0, alpha=1.0, nc=128, mask="left_columns", width=3,
          chains=256, envelope=12, base_seed=2026101401)


C5 = dict(N1=10, N2=10, beta=10.0, U=2.0, mu=1.0, alpha=1.0, nc=192, mask="left_columns", width=5,
          configs=1024, order_min=2, order_max=14, base_seed=2026101401, tag=0xC5)


WORKLOADS = {"c2": C2, "c3": C3, "c4": C4, "c5": C5}


def configs(name):
    """(orders, vertex arrays) of a workload's seeded configuration set: C2 the
    8192 configurations (orders 1..10), C4 the 256 chain envelopes (orders
    2..12), C5 the 1024 configurations (orders 2..14)."""
    import rng
    if name == "c2":
        return rng.synthetic_configs(C2["base_seed"], C2["tag"], C2["configs"])
    if name == "c4":
        return rng.synthetic_configs(C4["base_seed"], 0xC4, C4["chains"], 2, C4["envelope"], squared=True)
    if name == "c5":
        return rng.synthetic_configs(C5["base_seed"], C5["tag"], C5["configs"], C5["order_min"], C5["order_max"],
                                     squared=True)
    raise ValueError(f"no configuration set for workload {name}")


def context(name, derived=None):
    """A GPU context of the workload's lattice with its tables built on the
    device (C5 in uneven node chunks of 7 Nc + 3)."""
    import cdet_amd
    w = WORKLOADS[name]
    p = cdet_amd.make_params(w["N1"], w["N2"], w["nc"], w["beta"], w["U"], w["mu"], w["alpha"], w["alpha"],
                             derived=derived)
    ctx = cdet_amd.Context(p)
    region = w.get("radius", w.get("width", 0))
    ctx.build_tables(pbc=True, mask=w["mask"], mask_param=region,
                     node_chunk=7 * w["nc"] + 3 if name == "c5" else 0)
    return ctx

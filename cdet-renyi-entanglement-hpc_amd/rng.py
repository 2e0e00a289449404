"""Host random streams of the reference driver, vectorised over many engines.

* mt19937_64 (Matsumoto & Nishimura's 64-bit Mersenne Twister, standard
  parameters) with libstdc++'s uniform_real_distribution<double>:
  u = double(g()) / 2^64 (nudged below 1), x = u * (b - a) + a.
* derive_mcmcmc_seed: splitmix64 mixing of (base seed, stream tag, index)
  (include/cdet/mcmcmc_sampler.h:159-172).

Used to draw the seeded synthetic workloads (one engine per configuration or
chain) exactly as a C++ host would.
"""
import numpy as np

_M64 = (1 << 64) - 1
_N, _M = 312, 156
_MATRIX_A = np.uint64(0xB5026F5AA96619E9)
_UPPER = np.uint64(0xFFFFFFFF80000000)
_LOWER = np.uint64(0x7FFFFFFF)


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def derive_mcmcmc_seed(base_seed, stream_tag, sequence_index=0):
    mixed = splitmix64((base_seed ^ stream_tag) & _M64)
    mixed ^= (0x9E3779B97F4A7C15 * (sequence_index + 1)) & _M64
    return splitmix64(mixed & _M64)


class MT19937_64:
    """`count` independent mt19937_64 engines stepped in lockstep."""

    def __init__(self, seeds):
        seeds = np.asarray(seeds, dtype=np.uint64).reshape(-1)
        k = len(seeds)
        mt = np.zeros((k, _N), dtype=np.uint64)
        mt[:, 0] = seeds
        f = np.uint64(6364136223846793005)
        with np.errstate(over="ignore"):
            for i in range(1, _N):
                prev = mt[:, i - 1]
                mt[:, i] = f * (prev ^ (prev >> np.uint64(62))) + np.uint64(i)
        self.mt = mt
        self.idx = _N

    def _twist(self):
        mt = self.mt
        for i in range(_N):
            x = (mt[:, i] & _UPPER) | (mt[:, (i + 1) % _N] & _LOWER)
            xa = x >> np.uint64(1)
            xa = np.where((x & np.uint64(1)) != 0, xa ^ _MATRIX_A, xa)
            mt[:, i] = mt[:, (i + _M) % _N] ^ xa
        self.idx = 0

    def next_u64(self):
        if self.idx >= _N:
            self._twist()
        y = self.mt[:, self.idx].copy()
        self.idx += 1
        y ^= (y >> np.uint64(29)) & np.uint64(0x5555555555555555)
        y ^= (y << np.uint64(17)) & np.uint64(0x71D67FFFEDA60000)
        y ^= (y << np.uint64(37)) & np.uint64(0xFFF7EEE000000000)
        y ^= y >> np.uint64(43)
        return y

    def uniform(self, a=0.0, b=1.0):
        g = self.next_u64().astype(np.float64)  # round to nearest, as the C++ conversion
        u = g / 18446744073709551616.0
        u = np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)
        return u * (b - a) + a


def synthetic_configs(base_seed, tag, count, order_min=1, order_max=10, squared=False):
    """Seeded configurations: configuration i draws from
    mt19937_64(derive_mcmcmc_seed(base_seed, tag, i)) a uniform u, its order
    order_min + floor((order_max - order_min + 1) * u) (u^2 in place of u when
    `squared`, which makes high orders rare), and then, per vertex,
    (tau, x, y, replica) ~ U[0,1)^4."""
    seeds = [derive_mcmcmc_seed(base_seed, tag, i) for i in range(count)]
    eng = MT19937_64(seeds)
    span = order_max - order_min + 1
    u = eng.uniform()
    orders = order_min + np.floor(span * (u * u if squared else u)).astype(np.int64)
    orders = np.minimum(orders, order_max)
    draws = np.stack([eng.uniform() for _ in range(4 * order_max)], axis=1)
    verts = [draws[i, :4 * orders[i]].reshape(orders[i], 4) for i in range(count)]
    return orders, verts

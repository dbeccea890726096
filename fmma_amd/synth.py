"""Synthetic workloads, bit-identical on every host.

SplitMix64 (Steele, Lea & Flood 2014) counter streams give uniform doubles in
[0,1) as (z >> 11) * 2^-53; normals come from Box-Muller on two streams.  Seeds:
1 positions, 2 weights, 3 error-sample indices (the benchmark's convention).
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, offset=0):
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, n, offset=0):
    """n doubles uniform in [0, 1)."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, n):
    u1 = uniform(seed, n)
    u2 = uniform(seed, n, offset=n)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def points(seed, n, dim, lo=0.0, hi=1.0):
    return (lo + (hi - lo) * uniform(seed, n * dim)).reshape(n, dim)


def sample_indices(seed, n, k):
    """k distinct indices of range(n) (seeded, without replacement)."""
    if k >= n:
        return np.arange(n)
    keys = splitmix64(seed, n)
    return np.sort(np.argpartition(keys, k)[:k])


def gaussian_mixture(seed, n, dim, components=16, sigma=0.05):
    """n points from an isotropic Gaussian mixture, centres uniform in the unit cube."""
    centres = points(seed + 100, components, dim)
    comp = (splitmix64(seed + 200, n) % np.uint64(components)).astype(np.int64)
    z = normal(seed + 300, n * dim).reshape(n, dim)
    return centres[comp] + sigma * z


def batch_counts(seed, nprob, lo=1024, hi=65536):
    """nprob point counts, log-uniform in [lo, hi]."""
    u = uniform(seed, nprob)
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64).clip(lo, hi)


def batch_points(seed, counts, dim):
    """Concatenated points, uniform in each problem's own unit cube."""
    return points(seed, int(np.sum(counts)), dim)


def sample2_inputs(sizes):
    """The inputs of the reference's 2-D benchmark (benchmark/src/sample2.cpp):
    glibc rand() from its default seed, for N = sizes[0], sizes[1], ... in
    turn; per point source[n][0], target[n][0], source[n][1], target[n][1],
    each rand()/RAND_MAX.  Returns {N: (source [N,2], target [N,2])}."""
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    libc.srand(1)
    out = {}
    for n in sizes:
        v = np.fromiter((libc.rand() for _ in range(4 * n)), dtype=np.float64, count=4 * n) / 2147483647.0
        v = v.reshape(n, 2, 2)  # [point][dim][source, target]
        out[n] = (np.ascontiguousarray(v[:, :, 0]), np.ascontiguousarray(v[:, :, 1]))
    return out


# The five benchmark workloads, as (name, description) -- parameters follow
# the task's configuration list; bench.py and the tests build them from these.
def config(name, n=None):
    """Inputs of a named workload; n overrides the point count."""
    if name == "c1_1d":
        n = n or 4096
        return dict(dim=1, poly_ord=7, depth=-1, kernel="inv_diff", params=(), trans_sym=1,
                    targets=points(1, n, 1), sources=points(4, n, 1), weights=2.0 * uniform(2, n) - 1.0)
    if name == "c2_2d":
        n = n or 1 << 20
        x = points(1, n, 2)
        return dict(dim=2, poly_ord=5, depth=8, kernel="log_r", params=(), trans_sym=1,
                    targets=x, sources=x, weights=normal(2, n))
    if name == "c3_3d":
        n = n or 1 << 22
        x = points(1, n, 3)
        return dict(dim=3, poly_ord=3, depth=-1, kernel="inv_r_reg", params=(), trans_sym=1,
                    targets=x, sources=x, weights=uniform(2, n))
    if name == "c4_gauss":
        n = n or 1 << 20
        return dict(dim=3, poly_ord=3, depth=-1, kernel="gauss", params=(0.05,), trans_sym=1,
                    targets=points(5, n, 3), sources=gaussian_mixture(1, n, 3), weights=normal(2, n))
    if name == "c5_yukawa":
        n = n or 4096
        x = points(1, n, 3)
        return dict(dim=3, poly_ord=4, depth=-1, kernel="yukawa", params=(2.0,), trans_sym=1,
                    targets=x, sources=x, weights=normal(2, n))
    if name == "c5_batch":
        nprob = n or 512
        counts = batch_counts(7, nprob)
        x = batch_points(8, counts, 3)
        return dict(dim=3, poly_ord=4, depth=-1, kernel="yukawa", params=(2.0,), trans_sym=1, counts=counts,
                    targets=x, sources=x, weights=normal(9, x.shape[0]))
    raise KeyError(name)

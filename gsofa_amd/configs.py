"""The benchmark workloads of this path, with fixed seeds.

grid64_natural   5-point 2D Laplacian pattern, 64 x 64 grid, natural row-major
                 order (n=4,096).
grid3d_nd        7-point 3D Laplacian, 48^3 grid, geometric nested dissection
                 (leaf 64); each row gains with probability 0.05 one extra
                 one-directional entry to a grid point at Manhattan distance 2
                 (n=110,592).
band             n=659,033 banded unsymmetric: diagonal + Poisson(8)
                 off-diagonals drawn uniformly from |i-j| <= 256, duplicates
                 removed (the original test scale).
grid2d_nd_batch  256 2D 5-point grids, side log-uniform in [16,256], nested
                 dissection, per-row Bernoulli(0.05) extra entry at distance
                 2..3.
stream           request stream: Poisson arrivals; each request is a matrix of
                 the grid2d_nd_batch family or a 3D 7-point nested-dissection
                 grid with side uniform in [16,64].
"""
import math

from . import _lib, gen_band, gen_grid

SEEDS = {"grid64": 1, "grid3d": 2, "band": 3, "batch": 4, "stream": 5}
ND_LEAF = 64


def grid64_natural():
    return gen_grid(64, 64, 1, nd_leaf=0, seed=SEEDS["grid64"])


def grid3d_nd(side=48):
    return gen_grid(side, side, side, nd_leaf=ND_LEAF, extra_prob=0.05, extra_dist=2,
                    seed=SEEDS["grid3d"])


def band(n=659033):
    return gen_band(n, 256, 8.0, SEEDS["band"])


def batch_sides(count=256, seed=SEEDS["batch"], lo=16, hi=256):
    L = _lib.lib()
    return [int(L.gsofa_gen_side_loguniform(seed, b, lo, hi)) for b in range(count)]


def batch_matrix(side, index, seed=SEEDS["batch"]):
    return gen_grid(side, side, 1, nd_leaf=ND_LEAF, extra_prob=0.05, extra_dist=3,
                    seed=seed * 1000003 + index)


def grid2d_nd_batch(count=256, seed=SEEDS["batch"], max_side=256):
    sides = batch_sides(count, seed, 16, max_side)
    return [batch_matrix(s, b, seed) for b, s in enumerate(sides)]


def stream(count, rate, seed=SEEDS["stream"]):
    """Deterministic request stream: list of (arrival_s, kind, side, matrix).

    kind 0 = grid2d_nd_batch-family 2D grid (side log-uniform [16,256]), kind 1 = 3D 7-point
    ND grid (side uniform [16,64]).  Inter-arrival times are exponential with
    the given rate (requests per second).
    """
    L = _lib.lib()
    t = 0.0
    out = []
    for i in range(count):
        u = L.gsofa_gen_unit(seed, 1, i)
        t += -math.log(1.0 - u) / rate
        kind = 0 if L.gsofa_gen_unit(seed, 2, i) < 0.5 else 1
        if kind == 0:
            side = int(L.gsofa_gen_side_loguniform(seed * 7919 + 1, i, 16, 256))
            m = batch_matrix(side, i, seed)
        else:
            side = int(L.gsofa_gen_uniform_int(seed, 3, i, 16, 64))
            m = gen_grid(side, side, side, nd_leaf=ND_LEAF, seed=seed * 1000003 + i)
        out.append((t, kind, side, m))
    return out

"""The sampling pool's prefix sums and search, emulated on the CPU in the kernels' FP64 order.

k_pool_tile / k_pool_super / k_pool_top (qgate_amd/csrc/k_getter.hip) build a three-level CDF:
chunks of 16 bins summed serially, the 256 chunk sums of a tile scanned by one wavefront (four
chunks per lane; the lanes' totals scanned with DPP moves: within rows of 16 at distances 1, 2,
4, 8, then row 0's total into row 1 and row 2's into row 3, then lane 31's into rows 2 and 3;
each lane's exclusive prefix, the scan shifted by one lane, added to its four partial sums),
tiles and super-tiles scanned serially; k_pool_sample searches it level by level. This
emulation repeats those operations in the same order (numpy float64 is IEEE binary64, as the
GPU's FP64), so it checks the invariants the kernels rely on, for zeros, sparse and dense
inputs and pools smaller than one tile:
  * the tile total stored for the upper levels equals the tile's last chunk prefix, and each
    super-tile prefix equals the previous one plus its last tile prefix, bit for bit;
  * every draw, including draws exactly at a stored prefix and one ulp either side, returns a
    bin with p > 0 below the pool size;
  * a draw returns the bin upper_bound gives on the exact (rational) CDF, except within
    1e-13 of a bin edge.
"""
import bisect
from fractions import Fraction

import numpy as np
import pytest

CHUNK, TILE_CHUNKS, SUPER_TILES = 16, 256, 1024
TILE = CHUNK * TILE_CHUNKS


def build(p):
    """e_inc, m_inc (per tile, after the super-tile scan), h_inc, tile totals, as the kernels."""
    n = len(p)
    n_tiles = (n + TILE - 1) // TILE
    pp = np.zeros(n_tiles * TILE)
    pp[:n] = p
    e_inc = np.zeros(n_tiles * TILE_CHUNKS)
    t_tot = np.zeros(n_tiles)
    for t in range(n_tiles):
        x = pp[t * TILE:(t + 1) * TILE].reshape(TILE_CHUNKS, CHUNK)
        sc = np.zeros(TILE_CHUNKS)
        for c in range(TILE_CHUNKS):
            s = x[c, 0]
            for k in range(1, CHUNK):
                s = s + x[c, k]
            sc[c] = s
        q = np.zeros((64, 4))
        for l in range(64):
            q[l, 0] = sc[4 * l]
            q[l, 1] = q[l, 0] + sc[4 * l + 1]
            q[l, 2] = q[l, 1] + sc[4 * l + 2]
            q[l, 3] = q[l, 2] + sc[4 * l + 3]
        incl = q[:, 3].copy()
        for d in (1, 2, 4, 8):          # DPP row_shr:d, within rows of 16 lanes
            prev = incl.copy()          # a DPP move reads the values before this step
            for l in range(64):
                if l % 16 >= d:
                    incl[l] = incl[l] + prev[l - d]
        prev = incl.copy()              # row_bcast:15 into rows 1 and 3
        for l in list(range(16, 32)) + list(range(48, 64)):
            incl[l] = incl[l] + prev[(l // 16) * 16 - 1]
        prev = incl.copy()              # row_bcast:31 into rows 2 and 3
        for l in range(32, 64):
            incl[l] = incl[l] + prev[31]
        before = np.concatenate([[0.], incl[:-1]])   # wave_shr:1
        for l in range(64):
            for j in range(4):
                e_inc[t * TILE_CHUNKS + 4 * l + j] = before[l] + q[l, j]
        t_tot[t] = before[63] + q[63, 3]
    n_supers = (n_tiles + SUPER_TILES - 1) // SUPER_TILES
    m_inc = np.zeros(n_tiles)
    s_tot = np.zeros(n_supers)
    for s in range(n_supers):
        acc = 0.
        for i in range(SUPER_TILES):
            v = t_tot[s * SUPER_TILES + i] if s * SUPER_TILES + i < n_tiles else 0.
            acc = v if i == 0 else acc + v
            if s * SUPER_TILES + i < n_tiles:
                m_inc[s * SUPER_TILES + i] = acc
        s_tot[s] = acc
    h_inc = np.zeros(n_supers)
    acc = 0.
    for i in range(n_supers):
        acc = s_tot[0] if i == 0 else acc + s_tot[i]
        h_inc[i] = acc
    return pp, e_inc, m_inc, h_inc, t_tot


def first_true(lo, hi, f):
    while lo < hi:
        mid = lo + (hi - lo) // 2
        if f(mid):
            hi = mid
        else:
            lo = mid + 1
    return lo


def sample(pool, x):
    pp, e_inc, m_inc, h_inc, t_last, s_last, last_nonzero, norm = pool
    n_tiles, n_supers = len(m_inc), len(h_inc)
    g = first_true(0, n_supers, lambda s: h_inc[s] * norm > x)
    if g == n_supers:
        return last_nonzero
    hs = h_inc[g - 1] if g > 0 else 0.
    t0, t1 = g * SUPER_TILES, min(g * SUPER_TILES + SUPER_TILES, n_tiles)
    t = first_true(t0, t1, lambda k: (hs + m_inc[k]) * norm > x)
    if t == t1:
        return s_last[g]
    ms = m_inc[t - 1] if t > t0 else 0.
    c0, c1 = t * TILE_CHUNKS, t * TILE_CHUNKS + TILE_CHUNKS
    c = first_true(c0, c1, lambda k: (hs + (ms + e_inc[k])) * norm > x)
    if c == c1:
        return t_last[t]
    es = e_inc[c - 1] if c > c0 else 0.
    run, last = 0., -1
    for k in range(CHUNK):
        pk = pp[c * CHUNK + k]
        run = pk if k == 0 else run + pk
        if pk > 0.:
            last = c * CHUNK + k
        if (hs + (ms + (es + run))) * norm > x:
            return c * CHUNK + k
    return last if last >= 0 else last_nonzero


def make_pool(p):
    pp, e_inc, m_inc, h_inc, t_tot = build(p)
    nz = np.nonzero(pp > 0)[0]
    tiles = len(m_inc)
    t_last = [int(max([i for i in nz if i // TILE == t], default=-1)) for t in range(tiles)]
    s_last = [int(max([i for i in nz if i // (TILE * SUPER_TILES) == s], default=-1)) for s in range(len(h_inc))]
    total = h_inc[-1]
    return (pp, e_inc, m_inc, h_inc, t_last, s_last, int(nz[-1]), 1. / total), t_tot


def distributions():
    rs = np.random.RandomState(17)
    yield 'dense 2^13', rs.random_sample(1 << 13) ** 3
    p = np.zeros(1 << 14)
    p[rs.choice(1 << 14, 40, replace=False)] = rs.random_sample(40)
    yield 'sparse 2^14', p
    p = rs.random_sample(1 << 13)
    p[::2] = 0.
    p[4096:4096 + 600] = 0.
    yield 'zero runs 2^13', p
    p = np.zeros(1 << 12)
    p[[3, 4095]] = [0.25, 0.75]
    yield 'two bins 2^12', p
    yield 'tiny pool 2^3', np.array([0., 0.5, 0., 0., 0.125, 0., 0.375, 0.])
    yield 'uniform 2^12', np.full(1 << 12, 2. ** -12)
    yield 'scale 1e-300 2^12', rs.random_sample(1 << 12) * 1e-300


@pytest.mark.parametrize('name,p', list(distributions()), ids=[d[0] for d in distributions()])
def test_pool_scan_invariants_and_search(name, p):
    pool, t_tot = make_pool(p)
    pp, e_inc, m_inc, h_inc = pool[:4]
    n_tiles = len(m_inc)
    # levels agree exactly at the ends
    for t in range(n_tiles):
        assert t_tot[t] == e_inc[t * TILE_CHUNKS + TILE_CHUNKS - 1]
        s0 = (t // SUPER_TILES) * SUPER_TILES
        want = t_tot[t] if t == s0 else m_inc[t - 1] + t_tot[t]
        assert m_inc[t] == want
    norm = pool[7]
    # draws: random, at every stored prefix (all levels) and one ulp either side, and the ends
    xs = list(np.random.RandomState(3).random_sample(3000))
    prefixes = [v * norm for v in list(e_inc) + list(m_inc) + list(h_inc)]
    for v in prefixes:
        xs += [v, np.nextafter(v, 0.), np.nextafter(v, 2.)]
    xs += [0., np.nextafter(0., 1.), np.nextafter(1., 0.), 1., 1.5]
    xs = [x for x in xs if x >= 0.]
    # the exact CDF in rationals, for the upper_bound check
    exact = list(np.cumsum([Fraction(float(v)) for v in p], dtype=object))
    total = exact[-1]
    for x in xs:
        b = sample(pool, x)
        assert 0 <= b < len(p), (name, x, b)
        assert p[b] > 0., (name, x, b)
        if x >= 1.:
            continue
        # upper_bound on the exact CDF: the first bin whose cumulative probability exceeds x
        ub = bisect.bisect_right(exact, Fraction(float(x)) * total)
        if ub != b:
            edges = range(max(min(ub, b) - 1, 0), min(max(ub, b) + 1, len(p)))
            assert any(abs(x - float(exact[k] / total)) < 1e-13 for k in edges), (name, x, ub, b)

"""The resampling gather's partition (k_gather in csrc/pf_kernels.hip), restated in exact
integers and checked against brute-force systematic ancestors on skewed weights.

The gather takes items [2048 b, 2048 b + 2048) of the merged sequence of outputs (keys
K_i = K_0 + i T) and particles (cum_j n), output i ahead of particle j when K_i <= cum_j n.  It
finds the tiles holding J(d0) and J(d1) (the particles among the first d items) from the
per-tile end cumulative weights, stages the tiles from J(d0)'s on (at most three: the window
bound J(d1) <= J(d0) + 2048), and lets every lane merge eight items from a binary search of its
own diagonal.  These tests check that every output is written exactly once with its ancestor
(the first j with cum_j n >= K_i, pf_kernels.cu:660-676 in integer form), that every access
stays inside the staged window, and that the multi-round tile search (1024 probes per round in
the kernel; fewer here, to reach several rounds at small sizes) lands on the same tile.
"""
import random

import pytest

KT = 1024  # particles per tile
ITEMS = 2048  # merge-path items per block
WIN = 3 * KT  # staged window


def partition(q, U, probes):
    n = len(q)
    nt = (n + KT - 1) // KT
    cum, acc = [], 0
    for v in q + [0] * (nt * KT - n):
        acc += v
        cum.append(acc)
    T = acc
    assert T > 0
    tcum = [cum[t * KT + KT - 1] for t in range(nt)]
    K0 = -(-(T * U) // 2 ** 32)
    assert 0 <= K0 <= T

    def key(i):
        return K0 + i * T

    def cn(e):
        return cum[e] * n

    def among(e, d):  # particle e lies among the first d merged items
        i = d - 1 - e
        if i < 0:
            return False
        if i >= n:
            return True
        return cn(e) < key(i)

    rounds = 1 if nt <= probes else (2 if nt <= probes * probes else 3)
    anc = [None] * n
    for b in range(nt):
        d0 = min(b * ITEMS, 2 * n)
        d1 = min(d0 + ITEMS, 2 * n)
        if d0 >= d1:
            continue
        tiles = []
        for d in (d0, d1):
            lo, hi = 0, nt
            for _ in range(rounds):
                span = hi - lo
                step = (span + probes - 1) // probes
                if span == 0:
                    continue
                kf = None
                for k in range(probes):
                    if k * step >= span:
                        break
                    e = min((lo + k * step) * KT + KT - 1, n - 1)
                    if not among(e, d):
                        kf = k
                        break
                if kf is None:
                    lo += ((span - 1) // step) * step + 1
                else:
                    hi = lo + kf * step
                    lo = lo if kf == 0 else lo + (kf - 1) * step + 1
            assert lo == hi
            tiles.append(lo)
        ta, tb = tiles
        assert ta < nt and (tb == nt or tb <= ta + 2)
        wb = ta * KT
        wn = min(WIN, nt * KT - wb)
        js = []
        for d, tile in ((d0, ta), (d1, tb)):
            if tile >= nt:
                js.append(n)
                continue
            A, E = tile * KT, min(tile * KT + KT - 1, n - 1)
            assert not among(E, d) and (tile == 0 or among(A - 1, d))
            js.append(next(e for e in range(A, E + 1) if not among(e, d)))
        ja, jb = js
        ia, ib = d0 - ja, d1 - jb
        assert 0 <= ia <= ib <= n and jb - ja <= ITEMS
        for lane in range(256):
            d = d0 + 8 * lane
            if d >= d1:
                continue
            lo, hi = ja, jb
            while lo < hi:
                mid = (lo + hi) >> 1
                assert 0 <= mid - wb < wn
                if among(mid, d):
                    lo = mid + 1
                else:
                    hi = mid
            e, i = lo, d - lo
            for _ in range(d, min(d + 8, d1)):
                if i < n and (e >= n or not cn(e) < key(i)):
                    assert ia <= i < ib and anc[i] is None and 0 <= e - wb < wn
                    anc[i] = e
                    i += 1
                else:
                    assert 0 <= e - wb < wn
                    e += 1
    return anc, cum, K0, T


def brute(cum, n, K0, T):
    out, j = [], 0
    for i in range(n):
        while cum[j] * n < K0 + i * T:
            j += 1
        out.append(j)
    return out


def weights(rng, n, kind):
    if kind == "random":
        return [rng.randrange(1 << 20) for _ in range(n)]
    if kind == "one":
        q = [0] * n
        q[rng.randrange(n)] = 1 << 56
        return q
    if kind == "sparse":
        q = [(1 << 50) if rng.random() < 0.001 else 0 for _ in range(n)]
        q[-1] += 1
        return q
    if kind == "flat":
        return [1 << 40] * n
    if kind == "ends":
        q = [0] * n
        q[0], q[-1] = 5, 7
        return q
    if kind == "peaked":  # a few heavy particles among near-zero ones (the tracking regime)
        return [rng.randrange(1 << 56) if rng.random() < 0.01 else rng.randrange(3) for _ in range(n)]
    return [rng.randrange(1 << 56) for _ in range(n)]


KINDS = ["random", "one", "sparse", "flat", "ends", "peaked", "full"]


@pytest.mark.parametrize("kind", KINDS)
def test_partition_matches_brute_force(kind):
    rng = random.Random(hash(kind) & 0xFFFF)
    for n in (1, 2, 1023, 1024, 1025, 2049, 3000, 9001):
        for U in (0, 1, 2 ** 32 - 1, rng.randrange(2 ** 32)):
            q = weights(rng, n, kind)
            if sum(q) == 0:
                continue
            anc, cum, K0, T = partition(q, U, 1024)
            assert anc == brute(cum, n, K0, T), (kind, n, U)


@pytest.mark.parametrize("kind", ["random", "one", "peaked"])
def test_multi_round_tile_search(kind):
    """Four probes per round: filters of 5 to 64 tiles need two or three rounds."""
    rng = random.Random(7 + len(kind))
    for n in (4097, 17000, 40000, 65000):
        q = weights(rng, n, kind)
        anc, cum, K0, T = partition(q, rng.randrange(2 ** 32), 4)
        assert anc == brute(cum, n, K0, T), (kind, n)

"""Host-side checks of index arithmetic the kernels use in place of a direct
definition, on edge cases the GPU parity tests are unlikely to reach (a
uniform draw of 2^32 - 1, thresholds of 2^32 - 1, tiny tables, empty
segments). Each check compares the kernels' formulation, written out here,
with the definition it replaces."""
from __future__ import annotations

import numpy as np

U32 = 0xFFFFFFFF


def k_pick_direct(u, T, err_n):
    """Number of thresholds k < err_n - 1 with u >= T[k], capped at 4."""
    return min(4, sum(1 for k in range(max(err_n - 1, 0)) if u >= T[k]))


def k_pick_padded(u, T, err_n):
    """The kernels' form: four compares against the table padded with
    2^32 - 1 to four entries, capped at err_n - 1 (0 for an empty table)."""
    P = list(T[:err_n]) + [U32] * max(0, 4 - err_n)
    k = sum(1 for q in range(4) if u >= P[q])
    return min(k, err_n - 1 if err_n else 0)


def test_error_count_pick_padding_is_exact():
    rng = np.random.default_rng(1)
    for err_n in range(0, 9):
        for _ in range(300):
            # a non-decreasing table whose last entry is 2^32 - 1 (the compact form)
            body = np.sort(rng.integers(0, 1 << 32, size=max(err_n - 1, 0), dtype=np.uint64)).tolist()
            if rng.random() < 0.3 and body:
                body[-1] = U32  # a kept threshold may itself be 2^32 - 1
            T = body + ([U32] if err_n else [])
            for u in (0, 1, U32, U32 - 1, int(rng.integers(0, 1 << 32)), *(T[:3])):
                assert k_pick_padded(u, T, err_n) == k_pick_direct(u, T, err_n), (err_n, u, T)


def cdf_pick(T, u):
    """First k with u < T[k] (T non-decreasing, last entry 2^32)."""
    for k, t in enumerate(T):
        if u < t:
            return k
    return len(T) - 1


def guide_table(T32, bits):
    g = []
    k = 0
    for i in range(1 << bits):
        target = i << (32 - bits)
        while k + 1 < len(T32) and T32[k] <= target:
            k += 1
        g.append(k)
    return g


def guided_pick_coarse(T32, guide_fine, fine_bits, gb, u):
    """The kernels' pick with a guide of 2^gb entries taken from the host's
    2^fine_bits guide (every 2^(fine_bits - gb)-th entry), binary search
    inside the bucket."""
    guide = guide_fine[:: 1 << (fine_bits - gb)]
    g = u >> (32 - gb)
    lo = guide[g]
    hi = guide[g + 1] if g < (1 << gb) - 1 else len(T32) - 1
    while lo < hi:
        mid = (lo + hi) >> 1
        if u < T32[mid]:
            hi = mid
        else:
            lo = mid + 1
    return lo


def test_coarser_guides_pick_the_same_entry():
    rng = np.random.default_rng(2)
    for n in (1, 2, 3, 17, 120, 600):
        p = rng.dirichlet(np.ones(n) * 0.3)
        cdf = np.minimum(np.round(np.cumsum(p) * 2.0 ** 32), 2.0 ** 32).astype(np.uint64)
        cdf[-1] = 1 << 32
        T = [int(x) for x in cdf]
        T32 = [min(t, U32) for t in T]
        fine = guide_table(T32, 11)
        for gb in (8, 10, 11):
            for u in [0, U32, *rng.integers(0, 1 << 32, size=400).tolist(), *[t for t in T32[:-1]]]:
                u = int(u)
                assert guided_pick_coarse(T32, fine, 11, gb, u) == cdf_pick(T, u), (n, gb, u)


def test_presence_by_leaf_counts_equals_search():
    """count(hi) > count(lo) over 64-rank blocks of membership bits equals
    'a sample leaf lies in [lo, hi)'."""
    rng = np.random.default_rng(3)
    for n_leaves in (1, 63, 64, 65, 200, 1000):
        for _ in range(20):
            k = int(rng.integers(1, n_leaves + 1))
            leaves = np.sort(rng.choice(n_leaves, size=k, replace=False))
            nb = n_leaves // 64 + 1
            bits = np.zeros(nb, np.uint64)
            for r in leaves:
                bits[r >> 6] |= np.uint64(1) << np.uint64(r & 63)
            prefix = np.concatenate([[0], np.cumsum([bin(int(b)).count("1") for b in bits])[:-1]])

            def below(x):
                b = int(bits[x >> 6]) & ((1 << (x & 63)) - 1)
                return int(prefix[x >> 6]) + bin(b).count("1")

            for _ in range(200):
                lo = int(rng.integers(0, n_leaves + 1))
                hi = int(rng.integers(lo, n_leaves + 1))
                direct = bool(((leaves >= lo) & (leaves < hi)).any())
                assert (below(hi) > below(lo)) == direct, (n_leaves, lo, hi)


def test_item_lookup_over_segments_with_empty_ones():
    """Item q belongs to the last segment whose first item is at or below q;
    segments without templates own no items and share their successor's
    first item."""
    rng = np.random.default_rng(4)
    for _ in range(200):
        n_seg = int(rng.integers(1, 12))
        spi = int(rng.integers(1, 5))
        strata = rng.integers(1, 9, size=n_seg)
        empty = rng.random(n_seg) < 0.3
        first, n = [], 0
        for s in range(n_seg):
            first.append(n)
            if not empty[s]:
                n += (int(strata[s]) + spi - 1) // spi
        expected = [(s, j0) for s in range(n_seg) if not empty[s] for j0 in range(0, int(strata[s]), spi)]
        assert len(expected) == n
        for q in range(n):
            a, z = 0, n_seg
            while a < z:
                mid = (a + z) >> 1
                if first[mid] <= q:
                    a = mid + 1
                else:
                    z = mid
            s = a - 1
            assert (s, (q - first[s]) * spi) == expected[q]

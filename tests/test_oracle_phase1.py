"""Pins the oracle's super-k-mer split, flags and buckets by the definitions
they must satisfy (bucket_ingester.hpp:36-127, super_kmer.hpp:85-97,246-256),
using the window minima pinned in test_oracle_primitives.py:

* consecutive supers of a run share exactly one k-mer (k-base overlap), the
  first starts at the run start and the last ends at the run end;
* every (k-1)-mer inside a super has the super's minimizer, the ones just
  outside differ (maximality);
* every k-mer of a run is owned by exactly one super, the one whose minimizer
  is min(lmin, rmin) of that k-mer (BCALM2 rule);
* BEGIN only on the first super of a run, END only on the last;
* bucket = (minimizer >> 1) % B.
"""
import numpy as np
import pytest

import oracle as O

rng = np.random.default_rng(7)
ACGT = np.frombuffer(b"ACGT", np.uint8)


def _check_run(codes, sup, k, m, B):
    L = len(codes)
    K = L - k + 1
    mn = O.window_min(codes, k, m)  # (k-1)-mers 0..K
    n = len(sup["start"])
    assert n >= 1
    st = sup["start"].astype(np.int64)
    ln = sup["len"].astype(np.int64)
    assert st[0] == 0 and st[-1] + ln[-1] == L
    assert np.all(ln >= k)
    assert np.all(st[1:] == st[:-1] + ln[:-1] - k)
    owner_count = np.zeros(K, np.int64)
    for i in range(n):
        a_k = st[i]                 # first k-mer index
        b_k = st[i] + ln[i] - k     # last k-mer index
        fl = int(sup["flags"][i])
        assert bool(fl & 1) == (i == 0)
        assert bool(fl & 2) == (i == n - 1)
        # (k-1)-mers carried by this super: a_k+1 .. b_k (+ a_k when first, + b_k+1 when last)
        lo = a_k if i == 0 else a_k + 1
        hi = b_k + 1 if i == n - 1 else b_k
        M = mn[lo]
        assert np.all(mn[lo:hi + 1] == M)
        if lo > 0:
            assert mn[lo - 1] != M
        if hi < K:
            assert mn[hi + 1] != M
        assert int(sup["bucket"][i]) == (int(M) >> 1) % B
        for x in range(a_k, b_k + 1):
            first, last = x == a_k, x == b_k
            owned = not ((first and not fl & 4) or (last and not fl & 8))
            if owned:
                owner_count[x] += 1
                assert min(mn[x], mn[x + 1]) == M
    assert np.all(owner_count == 1)


@pytest.mark.parametrize("k,m,B", [(31, 19, 64), (31, 21, 1024), (31, 12, 37), (63, 27, 4096),
                                   (15, 9, 5), (5, 3, 1)])
def test_supers_definitions(k, m, B):
    runs = [rng.integers(0, 4, 700, dtype=np.uint8),
            np.tile(rng.integers(0, 4, 9, dtype=np.uint8), 40),
            np.zeros(k + 40, np.uint8),
            rng.integers(0, 4, k, dtype=np.uint8),        # exactly one k-mer
            rng.integers(0, 4, k + 1, dtype=np.uint8)]
    for codes in runs:
        seq = ACGT[codes].tobytes()
        sup = O.supers(seq, k, m, B)
        _check_run(codes, sup, k, m, B)


def test_supers_split_on_non_acgt_and_case():
    k, m, B = 21, 11, 16
    a = rng.integers(0, 4, 200, dtype=np.uint8)
    b = rng.integers(0, 4, 300, dtype=np.uint8)
    short = (b"ACGT" * 8)[:k - 1]
    seq = (ACGT[a].tobytes().lower() + b"N\r" + short + b"x" + ACGT[b].tobytes())
    sup = O.supers(seq, k, m, B)
    # run 1 = a (lower case accepted), the short run between separators is
    # dropped, run 2 = b starting after the 'x'
    off2 = 200 + 2 + (k - 1) + 1
    first = sup["start"] < off2
    s1 = {kk: v[first] for kk, v in sup.items()}
    s2 = {kk: v[~first] for kk, v in sup.items()}
    s2["start"] = s2["start"] - off2
    _check_run(a, s1, k, m, B)
    _check_run(b, s2, k, m, B)


def test_short_runs_are_dropped():
    seq = b"ACGTACGTAC" + b"N" + b"ACGT"
    assert len(O.supers(seq, 11, 5, 4)["start"]) == 0

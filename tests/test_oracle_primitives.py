"""Pins the oracle's primitives.

* Against the reference's own header code compiled into oracle/_ref
  (kmer.hpp, phase1_bucket_write/minimizer.hpp, super_kmer.hpp): skipped where
  the reference tree is not mounted.
* Against mathematics that fixes them independently: the ntHash rolling
  recurrence equals the from-scratch definition, the canonical hash is
  invariant under reverse complement, reverse complement is an involution,
  a window minimum equals numpy's brute-force minimum.
"""
import numpy as np
import pytest

import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference tree not mounted")
rng = np.random.default_rng(12345)


def _runs():
    yield rng.integers(0, 4, 500, dtype=np.uint8)
    yield np.zeros(200, np.uint8)                      # homopolymer: every hash ties
    yield np.tile(np.array([0, 1], np.uint8), 150)     # period-2 repeat
    yield np.tile(rng.integers(0, 4, 7, dtype=np.uint8), 60)  # period-7 repeat
    yield np.concatenate([rng.integers(0, 4, 100, dtype=np.uint8), np.full(80, 3, np.uint8),
                          rng.integers(0, 4, 100, dtype=np.uint8)])


def test_best_m_known_answers():
    # README's 661k run used k=31 with the default m=12; compute_best_m(k) <= k.
    assert O.orc().orc_best_m(31) == 12
    for k in range(1, 80):
        assert 1 <= O.orc().orc_best_m(k) <= max(k, 1)


@needs_ref
def test_best_m_matches_reference():
    for k in range(1, 80):
        assert O.orc().orc_best_m(k) == O.ref().ref_compute_best_m(k)


@pytest.mark.parametrize("m", [2, 5, 12, 19, 21, 27, 31])
def test_nthash_canonical_symmetry(m):
    codes = rng.integers(0, 4, 400, dtype=np.uint8)
    h = O.mmer_hashes(codes, m)
    rc = (3 - codes[::-1]).astype(np.uint8)
    hr = O.mmer_hashes(rc, m)
    # the m-mer at p in the forward strand is the RC of the m-mer at n-m-p
    assert np.array_equal(h, hr[::-1])


@needs_ref
@pytest.mark.parametrize("m", [2, 5, 12, 19, 21, 27, 31])
def test_nthash_matches_reference_roll(m):
    for codes in _runs():
        h = O.mmer_hashes(codes, m)
        r = np.zeros_like(h)
        O.ref().ref_rolled_hashes(O._p(codes, O.U8P), len(codes), m, O._p(r, O.U64P))
        assert np.array_equal(h, r)


@pytest.mark.parametrize("k,m", [(31, 19), (31, 21), (31, 12), (63, 27), (15, 9), (5, 3)])
def test_window_min_bruteforce(k, m):
    for codes in _runs():
        w = O.window_min(codes, k, m)
        h = O.mmer_hashes(codes, m)
        W = k - m
        brute = np.array([h[j:j + W].min() for j in range(len(codes) - k + 2)], np.uint64)
        assert np.array_equal(w, brute)


@needs_ref
@pytest.mark.parametrize("k,m", [(31, 19), (31, 21), (31, 12), (63, 27), (15, 9), (5, 3)])
def test_window_min_matches_reference(k, m):
    for codes in _runs():
        w = O.window_min(codes, k, m)
        r = np.zeros_like(w)
        O.ref().ref_window_min(O._p(codes, O.U8P), len(codes), k, m, O._p(r, O.U64P))
        assert np.array_equal(w, r)


def _kmer_int(codes):
    x = 0
    for c in codes:
        x = (x << 2) | int(c)
    return x


@pytest.mark.parametrize("k", [1, 2, 15, 31])
def test_revcomp64(k):
    for _ in range(200):
        codes = rng.integers(0, 4, k, dtype=np.uint8)
        x = _kmer_int(codes)
        r = O.orc().orc_revcomp64(x, k)
        assert r == _kmer_int(3 - codes[::-1])
        assert O.orc().orc_revcomp64(r, k) == x


@needs_ref
@pytest.mark.parametrize("k", [1, 2, 15, 31])
def test_revcomp64_matches_reference(k):
    for _ in range(200):
        x = int(rng.integers(0, 1 << (2 * k), dtype=np.uint64))
        assert O.orc().orc_revcomp64(x, k) == O.ref().ref_revcomp64(x, k)


@pytest.mark.parametrize("k", [32, 47, 63])
def test_revcomp128(k):
    out = np.zeros(2, np.uint64)
    for _ in range(100):
        codes = rng.integers(0, 4, k, dtype=np.uint8)
        x = _kmer_int(codes)
        O.orc().orc_revcomp128(x & (2**64 - 1), x >> 64, k, O._p(out, O.U64P))
        got = int(out[0]) | (int(out[1]) << 64)
        assert got == _kmer_int(3 - codes[::-1])
        if O.ref_available():
            ref = np.zeros(2, np.uint64)
            O.ref().ref_revcomp128(x & (2**64 - 1), x >> 64, k, O._p(ref, O.U64P))
            assert np.array_equal(out, ref)


@needs_ref
def test_nuc_to_2bit_matches_reference():
    t = O.codes_of(bytes(range(256)))
    for c in range(256):
        assert O.ref().ref_nuc_to_2bit(bytes([c])) == t[c]

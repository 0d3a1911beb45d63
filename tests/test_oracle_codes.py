"""Pins the oracle's colour-set code and the bucket-record bytes.

Elias-gamma / delta (the code of the colour-set stream, hybrid_color_sets.hpp:41-96,
read back by stream.hpp:196-237 through jermp/bits' read_delta):
* code-word lengths equal the textbook ones for every x < 2^16 and at every
  power-of-two boundary up to 2^40 (|gamma(N)| = 2L + 1, |delta(N)| = L +
  2 floor(log2(L + 1)) + 1, N = x + 1, L = floor(log2 N)) -- independent of
  bit order;
* code-word bits equal the textbook code words (gamma: L zeros, then N in
  binary; delta: gamma(L + 1), then N without its leading 1) with every binary
  field written least significant bit first: jermp/bits' append_bits order.
  That LSB-first order is the one convention only the absent jermp/bits could
  pin; a swapped unary/binary order, a dropped terminator or an off-by-one in
  N fails here.
Hybrid sets: every size of set over several colour counts against a brute-force
encoder written from the three branch definitions (sparse gaps, n-bit bitmap,
complement gaps), thresholds floor(0.25 n) and floor(0.75 n) as the reference
computes them in floating point (streaming_color_set_dict.hpp:88-89).

Records: the reference's own varint_write / write_super_kmer_packed
(super_kmer.hpp:43-117, compiled into oracle/_ref) against a textbook LEB128
encoder and the independent record reader the GPU bucket-file tests use.
"""
import numpy as np
import pytest

import bucket_format as F
import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference tree not mounted")
rng = np.random.default_rng(2024)


def _bits_lsb(words, nbits):
    """The stream's first nbits bits, bit i of the stream = bit i % 64 of word i // 64."""
    return "".join(str((int(words[i >> 6]) >> (i & 63)) & 1) for i in range(nbits))


def _bin_msb(v, n):
    return format(v, "b").zfill(n)[-n:] if n else ""


def textbook_gamma(N):
    L = N.bit_length() - 1
    return "0" * L + _bin_msb(N, L + 1)


def textbook_delta(N):
    L = N.bit_length() - 1
    return textbook_gamma(L + 1) + _bin_msb(N, L)[-L:] if L else textbook_gamma(1)


def lsb_fields_gamma(N):
    """gamma(N) with its binary field (the L bits after the leading 1) LSB first."""
    L = N.bit_length() - 1
    return "0" * L + "1" + _bin_msb(N, L)[::-1]


def lsb_fields_delta(N):
    L = N.bit_length() - 1
    return lsb_fields_gamma(L + 1) + _bin_msb(N, L)[::-1]


def _delta(x):
    w = np.zeros(2, np.uint64)
    nb = np.zeros(1, np.uint64)
    O.orc().orc_write_delta(x, O._p(w, O.U64P), O._p(nb, O.U64P))
    return w, int(nb[0])


def _sample():
    xs = list(range(1 << 16))
    for e in range(1, 41):
        xs += [(1 << e) - 2, (1 << e) - 1, 1 << e]
    xs += [int(v) for v in rng.integers(0, 1 << 20, 2000)]
    return xs


def test_delta_lengths_are_textbook():
    for x in _sample():
        N = x + 1
        L = N.bit_length() - 1
        _, nb = _delta(x)
        assert nb == L + 2 * ((L + 1).bit_length() - 1) + 1 == len(textbook_delta(N)), x


def test_delta_bits_are_textbook_lsb_first():
    for x in _sample()[:: 7] + [0, 1, 2, 3, 6, 7, 8, 1 << 19, (1 << 20) - 1]:
        N = x + 1
        w, nb = _delta(x)
        got = _bits_lsb(w, nb)
        want = lsb_fields_delta(N)
        assert got == want, (x, got, want)
        # the unary part and its terminator are convention-free
        tb = textbook_delta(N)
        z = tb.index("1")
        assert got[:z + 1] == tb[:z + 1]


def test_thresholds_match_float_rule():
    n = np.arange(0, 1 << 20, dtype=np.uint64)
    assert np.array_equal((0.25 * n.astype(np.float64)).astype(np.uint64), n // 4)
    assert np.array_equal((0.75 * n.astype(np.float64)).astype(np.uint64), (3 * n) // 4)


def brute_encode(colors, n):
    """The hybrid code of a sorted colour list over n colours, as a bit string."""
    N = len(colors)
    out = lsb_fields_delta(N + 1)
    if N < n // 4:
        prev = None
        for c in colors:
            out += lsb_fields_delta((c if prev is None else c - prev - 1) + 1)
            prev = c
    elif N < (3 * n) // 4:
        s = set(colors)
        out += "".join("1" if c in s else "0" for c in range(n))
    else:
        s = set(colors)
        prev = None
        for c in range(n):
            if c in s:
                continue
            out += lsb_fields_delta((c if prev is None else c - prev - 1) + 1)
            prev = c
    return out


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 9, 63, 64, 65, 100, 129, 500])
def test_hybrid_sets_match_brute_force(n):
    sizes = range(1, n + 1) if n <= 65 else sorted({1, 2, n // 4 - 1, n // 4, n // 4 + 1, n // 2,
                                                     (3 * n) // 4 - 1, (3 * n) // 4, (3 * n) // 4 + 1,
                                                     n - 1, n} | {int(v) for v in rng.integers(1, n + 1, 12)})
    for N in sizes:
        for _ in range(3):
            c = np.sort(rng.choice(n, size=N, replace=False)).astype(np.uint32)
            words, nbits = O.encode_set(c, n)
            want = brute_encode([int(v) for v in c], n)
            assert nbits == len(want), (n, N)
            assert _bits_lsb(words, nbits) == want, (n, N)


def leb128(v):
    out = bytearray()
    while True:
        b = v & 0x7F
        v >>= 7
        out.append(b | (0x80 if v else 0))
        if not v:
            return bytes(out)


@needs_ref
def test_reference_varint_is_leb128():
    vals = [0, 1, 127, 128, 255, 256, 16383, 16384, (1 << 21) - 1, 1 << 21, (1 << 32) - 1, 1 << 32,
            (1 << 63) - 1, 1 << 63, (1 << 64) - 1] + [int(v) for v in rng.integers(0, 1 << 62, 500)]
    buf = np.zeros(16, np.uint8)
    for v in vals:
        n = O.ref().ref_varint(v, O._p(buf, O.U8P))
        assert bytes(buf[:n]) == leb128(v), v
        assert F.varint(bytes(buf[:n]), 0) == (v, n)


@needs_ref
def test_reference_records_read_back():
    out = np.zeros(4096, np.uint8)
    stream = b""
    want = []
    for i in range(400):
        ln = int(rng.integers(1, 200)) if i % 5 else [1, 2, 3, 4, 5][i // 5 % 5]
        codes = rng.integers(0, 4, ln, dtype=np.uint8)
        nc = int(rng.integers(1, 6))
        cols = np.sort(rng.choice(1 << 20 if i % 2 else 40, size=nc, replace=False)).astype(np.uint32)
        fl = int(rng.integers(0, 16))
        n = O.ref().ref_record(fl, O._p(cols, O.U32P), nc, O._p(codes, O.U8P), ln, O._p(out, O.U8P))
        rec = bytes(out[:n])
        # varint(n_colours), varint(first), deltas, varint(len << 4 | flags), packed bases
        head = leb128(nc) + leb128(int(cols[0])) + b"".join(leb128(int(cols[j] - cols[j - 1])) for j in range(1, nc))
        head += leb128((ln << 4) | fl)
        assert rec[:len(head)] == head and len(rec) == len(head) + (ln + 3) // 4
        stream += rec
        want.append((tuple(int(c) for c in cols), fl, bytes(F.ACGT[c] for c in codes)))
    assert F.records(stream) == want

"""Host-side I/O of the drop-in boundary (no GPU needed).

* The LZ4 block codec of the bucket-file frames, pinned against the system
  liblz4 (LZ4_decompress_safe / LZ4_compress_default, the calls of
  bucket_io.hpp:255-265,695-696) in both directions.
* The FASTA/FASTQ reader, pinned by known answers that follow kseq_read's
  rules (external/kseq.h:161-201) and by plain == gzip.
"""
import ctypes
import ctypes.util
import gzip
import os

import numpy as np
import pytest

from conftest import cdbg_gpu as G

rng = np.random.default_rng(21)


def _syslz4():
    for name in ("liblz4.so.1", ctypes.util.find_library("lz4")):
        if not name:
            continue
        try:
            L = ctypes.CDLL(name)
            L.LZ4_decompress_safe.restype = ctypes.c_int
            L.LZ4_decompress_safe.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
            L.LZ4_compress_default.restype = ctypes.c_int
            L.LZ4_compress_default.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
            return L
        except OSError:
            continue
    return None


SYS = _syslz4()
needs_sys = pytest.mark.skipif(SYS is None, reason="system liblz4 not present")


def _inputs():
    yield b""
    yield b"A"
    yield b"ACGT" * 3
    yield bytes(rng.integers(0, 256, 5000, dtype=np.uint8))            # incompressible
    yield b"ACGTTGCA" * 20000                                          # long matches
    yield bytes(rng.integers(0, 4, 300000, dtype=np.uint8))            # 2-bit-ish records
    yield (b"\x01\x05" + bytes(rng.integers(0, 256, 9, dtype=np.uint8))) * 5000


def test_lz4_roundtrip_own():
    for src in _inputs():
        assert G.lz4_decompress(G.lz4_compress(src), len(src)) == src


@needs_sys
def test_lz4_own_blocks_decode_with_system_liblz4():
    for src in _inputs():
        c = G.lz4_compress(src)
        out = ctypes.create_string_buffer(max(len(src), 1))
        n = SYS.LZ4_decompress_safe(c, out, len(c), len(src))
        assert n == len(src)
        assert out.raw[:n] == src


@needs_sys
def test_lz4_system_blocks_decode_with_own():
    for src in _inputs():
        cap = len(src) + len(src) // 255 + 64
        buf = ctypes.create_string_buffer(cap)
        n = SYS.LZ4_compress_default(src, buf, len(src), cap)
        assert n > 0 or len(src) == 0
        assert G.lz4_decompress(buf.raw[:n], len(src)) == src


def _load(tmp_path, texts, gz=False):
    paths = []
    for i, t in enumerate(texts):
        p = str(tmp_path / (f"f{i}.fa" + (".gz" if gz else "")))
        with (gzip.open(p, "wb") if gz else open(p, "wb")) as f:
            f.write(t)
        paths.append(p)
    data, off = G.load_fasta(paths)
    return [bytes(data[off[i]:off[i + 1]]) for i in range(len(texts))]


KNOWN = [
    # (file content, expected colour bytes: each record followed by '\n')
    (b">r1 desc\nACGT\nacgt\n>r2\nNNA\n", b"ACGTacgt\nNNA\n"),
    (b"junk before\n>r\nAC\r\nGT\n", b"AC\rGT\n"),             # '\r' is kept
    (b">r\n\nAC\n\nGT\n", b"ACGT\n"),                           # empty lines skipped
    (b">a\nAAAA\n>b\n>c\nCC", b"AAAA\n\nCC\n"),                 # empty record, no final newline
    (b"@q1\nACGT\n+\nIIII\n@q2\nGG\n+q2\n!!\n", b"ACGT\nGG\n"),  # FASTQ
    (b"@q1\nACGT\n+\nIII", b""),                                # truncated quality drops the record
    (b">r\nAC\n+x\nGT\n", b"AC\n"),                              # '+' line: "GT" is AC's quality
    (b"", b""),
]


@pytest.mark.parametrize("text,want", KNOWN)
def test_fasta_known_answers(tmp_path, text, want):
    assert _load(tmp_path, [text]) == [want]


def test_fasta_plain_equals_gzip_and_colour_order(tmp_path):
    texts = []
    for i in range(3):
        seqs = [bytes(np.frombuffer(b"ACGTN", np.uint8)[rng.integers(0, 5, 500)]) for _ in range(4)]
        texts.append(b"".join(b">s%d\n" % j + b"\n".join(s[x:x + 60] for x in range(0, len(s), 60)) + b"\n"
                              for j, s in enumerate(seqs)))
    plain = _load(tmp_path, texts)
    (tmp_path / "gz").mkdir()
    gzd = _load(tmp_path / "gz", texts, gz=True)
    assert plain == gzd
    for t, got in zip(texts, plain):
        recs = [b"".join(r.split(b"\n")[1:]) for r in t.split(b">")[1:]]
        assert got == b"".join(r + b"\n" for r in recs)


def test_fasta_missing_file_is_an_error(tmp_path):
    with pytest.raises(RuntimeError):
        G.load_fasta([str(tmp_path / "nope.fa")])

"""Input decode on the device (cdbg_gpu_job_create_from_files).

The job's input bytes and colour offsets must equal the host reader's
(csrc/fasta.cpp, pinned by kseq known answers in test_host_io.py) on the same
files, plain and gzip: the known-answer texts, leading bytes before the first
record, '>' inside sequence lines, '@' headers, empty records and lines, no
closing newline, CRLF, files of several gzip members, gzip header fields,
every zlib level and strategy, and a multi-colour synthetic pangenome.  FASTQ
input goes to the host reader.  Corrupt gzip data is an error, not a short
read.
"""
import gzip
import struct
import zlib

import numpy as np
import pytest

from conftest import cdbg_gpu as G

pytestmark = pytest.mark.gpu

TEXTS = [
    b">r1 desc\nACGT\nacgt\n>r2\nNNA\n",
    b"junk before\n>r\nAC\r\nGT\n",
    b">r\n\nAC\n\nGT\n",
    b">a\nAAAA\n>b\n>c\nCC",
    b"xx>mid-line header\nACGT>kept\nGG\n",
    b"@h1\nACGT\n@h2\nTT\n",
    b">only header",
    b"no record at all\n",
    b"",
    b">r\r\nAC\r\nGT\r\n>s\r\n\r\nTT\r\n",
    b"\n\n>late\nA\nC\nG\nT\n\n\n",
]


def _write(tmp_path, name, data):
    p = str(tmp_path / name)
    with open(p, "wb") as f:
        f.write(data)
    return p


def _both(ctx, files, k=11, m=7, B=8):
    """The host reader's input and the job's, with gzip inflated on the device
    (gzip_device=1) and on host threads (2); both must agree."""
    data, off = G.load_fasta(files)
    out = []
    try:
        for mode in (1, 2):
            ctx.set_tuning(gzip_device=mode)
            with ctx.job_from_files(files, k, m, B) as job:
                got, on_dev = job.input()
                goff = job.input_offsets(len(files))
            out.append(((bytes(got), list(goff)), on_dev))
    finally:
        ctx.set_tuning()
    assert out[0] == out[1]
    return (bytes(data), list(off)), out[0][0], out[0][1]


def test_known_texts_plain_and_gzip(gpu_ctx, tmp_path):
    for gz in (False, True):
        files = [_write(tmp_path, f"t{i}{'.gz' if gz else ''}", gzip.compress(t) if gz else t)
                 for i, t in enumerate(TEXTS)]
        host, dev, on_dev = _both(gpu_ctx, files)
        assert on_dev
        assert dev == host, gz


def test_gzip_variants(gpu_ctx, tmp_path):
    rng = np.random.default_rng(9)
    seqs = [bytes(np.frombuffer(b"ACGTNacgt", np.uint8)[rng.integers(0, 9, n)]) for n in (0, 1, 70, 5000, 200000)]
    text = b"".join(b">s%d\n" % i + b"\n".join(s[x:x + 61] for x in range(0, len(s), 61)) + b"\n"
                    for i, s in enumerate(seqs))
    blobs = [gzip.compress(text, compresslevel=lvl) for lvl in (0, 1, 6, 9)]
    for strat in (zlib.Z_FIXED, zlib.Z_HUFFMAN_ONLY, zlib.Z_RLE):
        co = zlib.compressobj(6, zlib.DEFLATED, 31, 9, strat)
        blobs.append(co.compress(text) + co.flush())
    half = len(text) // 2
    blobs.append(gzip.compress(text[:half]) + gzip.compress(text[half:], 9))  # two members
    blobs.append(b"".join(gzip.compress(text[x:x + 4096]) for x in range(0, len(text), 4096)))  # many, as bgzip
    body = zlib.compressobj(6, zlib.DEFLATED, -15)
    raw = body.compress(text) + body.flush()
    hdr = b"\x1f\x8b\x08" + bytes([4 | 8 | 16 | 2]) + b"\0\0\0\0\0\xff" + struct.pack("<H", 3) + b"abc"
    hdr += b"x.fa\0note\0"
    blobs.append(hdr + struct.pack("<H", zlib.crc32(hdr) & 0xFFFF) + raw +
                 struct.pack("<II", zlib.crc32(text), len(text) & 0xFFFFFFFF))
    files = [_write(tmp_path, f"v{i}.fa.gz", b) for i, b in enumerate(blobs)]
    host, dev, on_dev = _both(gpu_ctx, files)
    assert on_dev and dev == host


def test_pangenome_files(gpu_ctx, tmp_path):
    paths = G.synth_fasta_files(str(tmp_path / "pg"), 12, 300000, gz=True, threads=4)
    host, dev, on_dev = _both(gpu_ctx, paths, 31, 19, 64)
    assert on_dev and dev == host


def test_fastq_goes_to_the_host_reader(gpu_ctx, tmp_path):
    files = [_write(tmp_path, "a.fa", b">a\nACGT\n"), _write(tmp_path, "q.fq.gz", gzip.compress(b"@q\nACGT\n+\nIIII\n"))]
    host, dev, on_dev = _both(gpu_ctx, files)
    assert not on_dev and dev == host


def test_corrupt_gzip_is_an_error(gpu_ctx, tmp_path):
    g = gzip.compress(b">r\n" + b"ACGT" * 5000 + b"\n")
    bad_crc = g[:-8] + struct.pack("<I", zlib.crc32(b"x")) + g[-4:]
    try:
        for mode in (1, 2):
            gpu_ctx.set_tuning(gzip_device=mode)
            for name, blob in (("crc.gz", bad_crc), ("cut.gz", g[:len(g) // 2] + g[-8:])):
                p = _write(tmp_path, name, blob)
                with pytest.raises(RuntimeError, match="gzip"):
                    gpu_ctx.job_from_files([p], 11, 7, 8)
    finally:
        gpu_ctx.set_tuning()

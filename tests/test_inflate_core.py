"""The device inflater's decoding core (csrc/inflate_core.h) on the host.

tests/native/inflate_host_check.cpp runs the same Inflater the device kernel
runs (one lane, a plain output) and is pinned here against zlib, the library
behind the reference's gzread (seq_reader.hpp:44-77): every stream zlib
writes (levels 0-9, fixed-Huffman, Huffman-only, RLE and filtered
strategies, stored blocks above 64 KiB, distances up to 32 KiB, runs needing
length bytes), gzip members with every optional header field, several
members in one file, trailing bytes after the last member, plain (not gzip)
input; each member's CRC-32 is checked with a bitwise CRC and the
combination rule the device uses.  Truncated streams and wrong trailers are
errors.
"""
import gzip
import os
import random
import shutil
import struct
import subprocess
import zlib

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def ihc(tmp_path_factory):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("ihc") / "ihc")
    subprocess.run([cxx, "-O2", "-std=c++17", "-Wall", "-Werror", "-o", exe,
                    os.path.join(HERE, "native", "inflate_host_check.cpp")], check=True)
    return exe


def _run(exe, tmp_path, data):
    src, dst = str(tmp_path / "in.gz"), str(tmp_path / "out")
    with open(src, "wb") as f:
        f.write(data)
    r = subprocess.run([exe, src, dst], capture_output=True, text=True)
    return r.returncode, (open(dst, "rb").read() if r.returncode == 0 else None)


def _sources():
    rnd = random.Random(5)
    dna = bytes(rnd.choice(b"ACGT") for _ in range(300000))
    fasta = b"".join(b">r%d\n" % i + bytes(rnd.choice(b"ACGTN") for _ in range(rnd.randint(0, 200))) + b"\n"
                     for i in range(3000))
    return {"empty": b"", "one": b"A", "dna100": dna[:100], "dna": dna, "fasta": fasta, "zeros": bytes(200000),
            "random": bytes(rnd.getrandbits(8) for _ in range(70000)), "ab": b"ab" * 50000}


@pytest.mark.parametrize("level", [0, 1, 6, 9])
def test_zlib_levels(ihc, tmp_path, level):
    for name, src in _sources().items():
        rc, out = _run(ihc, tmp_path, gzip.compress(src, compresslevel=level))
        assert rc == 0 and out == src, (level, name)


@pytest.mark.parametrize("strategy", [zlib.Z_FIXED, zlib.Z_HUFFMAN_ONLY, zlib.Z_RLE, zlib.Z_FILTERED])
def test_zlib_strategies(ihc, tmp_path, strategy):
    for name, src in _sources().items():
        co = zlib.compressobj(6, zlib.DEFLATED, 31, 9, strategy)
        rc, out = _run(ihc, tmp_path, co.compress(src) + co.flush())
        assert rc == 0 and out == src, (strategy, name)


def test_members_fields_garbage_plain(ihc, tmp_path):
    src = _sources()
    assert _run(ihc, tmp_path, gzip.compress(b"hello ") + gzip.compress(b"world", 9)) == (0, b"hello world")
    assert _run(ihc, tmp_path, gzip.compress(src["dna100"]) + b"xyzzy") == (0, src["dna100"])
    body = zlib.compressobj(6, zlib.DEFLATED, -15)
    raw = body.compress(src["fasta"][:5000]) + body.flush()
    hdr = b"\x1f\x8b\x08" + bytes([4 | 8 | 16 | 2]) + b"\0\0\0\0\0\xff" + struct.pack("<H", 5) + b"extra"
    hdr += b"name.fa\0comment\0"
    member = hdr + struct.pack("<H", zlib.crc32(hdr) & 0xFFFF) + raw
    member += struct.pack("<II", zlib.crc32(src["fasta"][:5000]), 5000)
    assert _run(ihc, tmp_path, member) == (0, src["fasta"][:5000])
    assert _run(ihc, tmp_path, src["fasta"][:2000]) == (0, src["fasta"][:2000])


def test_errors(ihc, tmp_path):
    g = gzip.compress(_sources()["dna"])
    for cut in (10, 11, 20, 100, len(g) // 2, len(g) - 9, len(g) - 1):
        assert _run(ihc, tmp_path, g[:cut])[0] != 0, cut
    assert _run(ihc, tmp_path, g[:-8] + struct.pack("<I", 1234) + g[-4:])[0] == 4  # CRC
    assert _run(ihc, tmp_path, g[:-4] + struct.pack("<I", 7))[0] == 3  # size

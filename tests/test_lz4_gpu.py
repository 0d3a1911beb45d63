"""The device LZ4 codec of the bucket-file frames (cdbg_gpu_lz4_*_frames).

Pinned against the system liblz4, the library behind the reference's
LZ4_compress_default / LZ4_decompress_safe calls (bucket_io.hpp:255-265,
695-696), in both directions: every device-compressed block must decode with
LZ4_decompress_safe to its input, and every block liblz4 makes (fast and HC
modes, whose matches, offsets and length bytes differ from the device's
greedy parse) must decode on the device.  Malformed blocks are rejected by the
decoder's own bounds checks.
"""
import ctypes

import numpy as np
import pytest

from test_host_io import SYS

pytestmark = pytest.mark.gpu
needs_sys = pytest.mark.skipif(SYS is None, reason="system liblz4 not present")

rng = np.random.default_rng(5)


def _frames():
    """Frames of 0..65536 bytes: empty, below the 13-byte match minimum,
    incompressible, long runs (length bytes past 255), 2-bit data, record-like
    bytes with short repeats, matches ending near the frame end."""
    out = [b"", b"A", b"ACGTACGTACGT", b"ACGTACGTACGTA", bytes(13), bytes(14)]
    out.append(bytes(rng.integers(0, 256, 5000, dtype=np.uint8)))
    out.append(bytes(65536))
    out.append(b"ACGTTGCA" * 8192)
    out.append(bytes(rng.integers(0, 4, 65536, dtype=np.uint8)))
    out.append((b"\x01\x05" + bytes(rng.integers(0, 256, 9, dtype=np.uint8))) * 5000)
    out.append(bytes(rng.integers(0, 256, 300, dtype=np.uint8)) * 3 + b"xyz")
    out.append(bytes(rng.integers(0, 256, 65536, dtype=np.uint8)))
    for n in (17, 100, 4095, 4096, 4097, 65535):
        base = bytes(rng.integers(0, 8, n, dtype=np.uint8))
        out.append(base)
    return out


def _sys_decode(blk, n):
    buf = ctypes.create_string_buffer(max(n, 1))
    got = SYS.LZ4_decompress_safe(blk, buf, len(blk), n)
    assert got == n
    return buf.raw[:n]


@needs_sys
def test_device_blocks_decode_with_system_liblz4(gpu_ctx):
    fr = _frames()
    blocks = gpu_ctx.lz4_compress_frames(fr)
    for f, b in zip(fr, blocks):
        assert len(b) <= len(f) + len(f) // 255 + 16
        assert _sys_decode(b, len(f)) == f
    # repetitive frames must actually compress
    assert len(blocks[7]) < 1024 and len(blocks[8]) < 1024


def test_device_roundtrip(gpu_ctx):
    fr = _frames()
    blocks = gpu_ctx.lz4_compress_frames(fr)
    assert gpu_ctx.lz4_decompress_frames(blocks, [len(f) for f in fr]) == fr


@needs_sys
@pytest.mark.parametrize("hc", [False, True])
def test_system_blocks_decode_on_device(gpu_ctx, hc):
    fr = _frames() + [b"GATTACA" * 40000]  # a frame above 64 KiB, as the reference's spills can be
    blocks = []
    for f in fr:
        cap = len(f) + len(f) // 255 + 64
        buf = ctypes.create_string_buffer(cap)
        n = SYS.LZ4_compress_HC(f, buf, len(f), cap, 9) if hc else SYS.LZ4_compress_default(f, buf, len(f), cap)
        assert n > 0 or len(f) == 0
        blocks.append(buf.raw[:n])
    assert gpu_ctx.lz4_decompress_frames(blocks, [len(f) for f in fr]) == fr


def test_malformed_blocks_rejected(gpu_ctx):
    good = gpu_ctx.lz4_compress_frames([b"ACGTTGCA" * 100])[0]
    cases = [
        (good[:-3], 800),                    # truncated
        (good, 799),                         # decodes past its slot
        (good, 801),                         # decodes short of its slot
        (bytes([0x0F, 1, 0]), 8),            # match with offset 1 before any output
        (bytes([0x10, 65, 0, 0]), 20),      # offset 0
        (bytes([0xF0, 255]), 300),           # literal length runs past the block
    ]
    for blk, n in cases:
        with pytest.raises(RuntimeError, match="malformed"):
            gpu_ctx.lz4_decompress_frames([blk], [n])


def test_frame_limit(gpu_ctx):
    with pytest.raises(RuntimeError, match="64 KiB"):
        gpu_ctx.lz4_compress_frames([bytes(65537)])

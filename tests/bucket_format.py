"""Independent reader of the bucket-file format, written from the reference's
format code: frames [u32 uncompressed][u32 compressed][LZ4 block] ... [u32 0]
(bucket_io.hpp:19-40) and records varint(n_colors), varint(first colour),
varint(colour deltas)..., varint(len << 4 | flags), ceil(len/4) bytes with
base i at bits 2(i%4) of byte i/4 (super_kmer.hpp:43-82,105-143)."""
import ctypes
import struct

ACGT = b"ACGT"


def varint(buf, pos):
    """LEB128: 7-bit groups, least significant first, high bit = more."""
    v = sh = 0
    while True:
        b = buf[pos]
        pos += 1
        v |= (b & 0x7F) << sh
        if not b & 0x80:
            return v, pos
        sh += 7


def frames(raw, decompress):
    """Concatenated payload of a bucket file's frames; decompress(block, u)
    returns the u bytes of one LZ4 block."""
    out, p = bytearray(), 0
    while p + 4 <= len(raw):
        u, = struct.unpack_from("<I", raw, p)
        p += 4
        if u == 0:
            break
        c, = struct.unpack_from("<I", raw, p)
        p += 4
        out += decompress(raw[p:p + c], u)
        p += c
    return bytes(out)


def records(d):
    """(colours, flags, ACGT bases) of every record of a decompressed stream;
    the padding bits of each record's last byte must be zero."""
    p, recs = 0, []
    while p < len(d):
        nc, p = varint(d, p)
        cols, prev = [], 0
        for _ in range(nc):
            x, p = varint(d, p)
            prev += x
            cols.append(prev)
        lf, p = varint(d, p)
        ln, fl = lf >> 4, lf & 15
        nb = (ln + 3) // 4
        bases = bytes(ACGT[(d[p + i // 4] >> (2 * (i % 4))) & 3] for i in range(ln))
        assert all((d[p + nb - 1] >> (2 * j)) & 3 == 0 for j in range(ln - 4 * (nb - 1), 4)), "padding bits"
        p += nb
        recs.append((tuple(cols), fl, bases))
    return recs


def sys_lz4_decompress(lib):
    """decompress() over a system liblz4 handle."""
    def dec(blk, u):
        buf = ctypes.create_string_buffer(u)
        assert lib.LZ4_decompress_safe(blk, buf, len(blk), u) == u
        return buf.raw
    return dec

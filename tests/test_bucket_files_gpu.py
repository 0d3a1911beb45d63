"""The drop-in file boundary on the GPU.

* cdbg_gpu_bucket_write: FASTA files -> bucket_<b>.bin.  Files are decoded
  with an independent reader (system liblz4 when present + a varint parser
  written from super_kmer.hpp:121-143) and must hold exactly the oracle's
  super-k-mer records merged as the compacting writer merges them
  (bucket_io.hpp:145-160,221-240): one record per distinct (bases,
  ownership bits) of a bucket, in order of first occurrence, with the sorted
  unique colours of its occurrences and the AND of their flags.
* cdbg_gpu_bucket_process: bucket files -> frag_unitigs.bin, .color_sets,
  .color_sets.tmp_offsets, byte-identical to the oracle's phase 2 over the
  same files.  Also on bucket files whose records carry several colours and
  arrive in another order (what the reference's compacting writer produces).
"""
import os
import struct

import numpy as np
import pytest

import bucket_format as F
import oracle as O
from conftest import cdbg_gpu as G
from test_host_io import SYS

pytestmark = pytest.mark.gpu


def _frames(path):
    raw = open(path, "rb").read()
    return F.frames(raw, F.sys_lz4_decompress(SYS) if SYS is not None else G.lz4_decompress)


def _records(path):
    return F.records(_frames(path))


def _fasta_job(tmp_path, colors=5, glen=20000, seed=6):
    data, off = G.synth(seed=seed, num_colors=colors, genome_len=glen, snp_rate=0.01, indel_rate=0.001)
    paths = []
    for c in range(colors):
        seq = bytes(data[off[c]:off[c + 1] - 1])
        p = str(tmp_path / f"c{c}.fa")
        with open(p, "wb") as f:
            f.write(b">g%d\n" % c + b"\n".join(seq[i:i + 70] for i in range(0, len(seq), 70)) + b"\n")
        paths.append(p)
    return paths


OWN = 0xC  # SK_FLAG_OWNS_FIRST | SK_FLAG_OWNS_LAST (super_kmer.hpp:96-97)


def _merged_bucket(data, sup, lo, hi):
    """The compacting writer's records for the oracle's super-k-mers
    [lo, hi) of one bucket (stream order): occurrences with equal bases and
    ownership bits merge; the record sits at its first occurrence, carries the
    sorted unique colours and the AND of the flags (bucket_io.hpp:145-160,
    221-240)."""
    groups = {}
    for i in range(lo, hi):
        s, ln = int(sup["start"][i]), int(sup["len"][i])
        fl = int(sup["flags"][i])
        key = (bytes(data[s:s + ln]).upper(), fl & OWN)
        g = groups.setdefault(key, [set(), 15])
        g[0].add(int(sup["color"][i]))
        g[1] &= fl
    return [(tuple(sorted(c)), fl, key[0]) for key, (c, fl) in groups.items()]


@pytest.mark.parametrize("k,m,B", [(31, 19, 64), (63, 27, 32), (21, 11, 7)])
def test_bucket_write_records(gpu_ctx, tmp_path, k, m, B):
    files = _fasta_job(tmp_path)
    d = str(tmp_path / "buckets")
    unc = gpu_ctx.bucket_write(files, k, m, B, d)
    data, off = G.load_fasta(files)
    ob = O.Build(data, off, k, m, B)
    sup = ob.supers(1)
    boff = np.concatenate([[0], np.cumsum(np.bincount(sup["bucket"], minlength=B))])
    merged = 0
    for b in range(B):
        path = os.path.join(d, f"bucket_{b}.bin")
        recs = _records(path)
        want = _merged_bucket(data, sup, boff[b], boff[b + 1])
        assert recs == want, f"bucket {b}"
        merged += sum(len(r[0]) > 1 for r in recs)
        payload = len(_frames(path))
        assert unc[b] == payload
        raw = open(path, "rb").read()
        p = 0
        while True:  # frames of at most 64 KiB, then the end marker
            u, = struct.unpack_from("<I", raw, p)
            if u == 0:
                assert p + 4 == len(raw)
                break
            c, = struct.unpack_from("<I", raw, p + 4)
            assert u <= 65536
            p += 8 + c
    assert merged > 0  # five related genomes share super-k-mers


def _compare_phase2(gpu_ctx, tmp_path, d, B, k, num_colors):
    frag, cs = str(tmp_path / "frag_unitigs.bin"), str(tmp_path / "out.color_sets")
    st = gpu_ctx.bucket_process(d, B, k, num_colors, frag, cs)
    ob = O.Build.from_bucket_files(d, B, k, num_colors)
    assert open(frag, "rb").read() == ob.frags().tobytes()
    words, offs = ob.color_sets()
    full = ob.stat("cs_bits") // 64
    assert open(cs, "rb").read() == bytes(28) + words[:full].tobytes()
    assert open(cs + ".tmp_offsets", "rb").read() == offs.tobytes()
    assert st["num_color_sets"] == ob.stat("num_sets")
    assert st["num_bits"] == ob.stat("cs_bits")
    assert st["num_kmers"] == ob.stat("num_kmers")
    assert st["trailing_word"] == (int(words[full]) if ob.stat("cs_bits") % 64 else 0)
    return st


@pytest.mark.parametrize("k,m,B", [(31, 19, 64), (63, 27, 32)])
def test_bucket_process_matches_oracle(gpu_ctx, tmp_path, k, m, B):
    files = _fasta_job(tmp_path, colors=6)
    d = str(tmp_path / "buckets")
    gpu_ctx.bucket_write(files, k, m, B, d)
    _compare_phase2(gpu_ctx, tmp_path, d, B, k, len(files))


@pytest.mark.parametrize("budget", [0, 1 << 20])
def test_bucket_process_ex(gpu_ctx, tmp_path, budget):
    """process_buckets' remaining parameters: the progress counter reaches
    the bucket count, a small device budget (many chunks) leaves the outputs
    unchanged, and consumed bucket files are deleted."""
    files = _fasta_job(tmp_path, colors=5)
    d = str(tmp_path / "buckets")
    B, k = 64, 31
    gpu_ctx.bucket_write(files, k, 19, B, d)
    ref = _compare_phase2(gpu_ctx, tmp_path, d, B, k, len(files))
    want = open(str(tmp_path / "frag_unitigs.bin"), "rb").read()
    frag, cs = str(tmp_path / "ex_frag.bin"), str(tmp_path / "ex.color_sets")
    st, done = gpu_ctx.bucket_process_ex(d, B, k, len(files), frag, cs, mem_budget_bytes=budget,
                                         delete_consumed=True)
    assert done == B
    assert st == ref
    assert open(frag, "rb").read() == want
    assert not any(os.path.exists(os.path.join(d, f"bucket_{b}.bin")) for b in range(B))


def test_file_path_equals_resident_path(gpu_ctx, tmp_path):
    files = _fasta_job(tmp_path, colors=4)
    data, off = G.load_fasta(files)
    with gpu_ctx.job(data, off, 31, 19, 64) as job:
        job.phase1()
        d = str(tmp_path / "b")
        job.write_bucket_files(d)
        job.phase2()
        resident = job.frags()
        rw, ro = job.color_sets()
    with gpu_ctx.job_from_bucket_files(d, 64, 31, 4) as fj:
        fj.run()
        assert np.array_equal(fj.frags(), resident)
        fw, fo = fj.color_sets()
        assert np.array_equal(fw, rw) and np.array_equal(fo, ro)


def _literal_block(b):
    n = len(b)
    tok = bytearray([min(n, 15) << 4])
    if n >= 15:
        r = n - 15
        while r >= 255:
            tok.append(255)
            r -= 255
        tok.append(r)
    return bytes(tok) + b


def _enc_varint(v):
    out = bytearray()
    while v >= 0x80:
        out.append((v & 0x7F) | 0x80)
        v >>= 7
    out.append(v)
    return bytes(out)


def _sys_block(blk):
    import ctypes
    cap = len(blk) + len(blk) // 255 + 64
    buf = ctypes.create_string_buffer(cap)
    n = SYS.LZ4_compress_default(blk, buf, len(blk), cap)
    assert n > 0
    return buf.raw[:n]


def _write_bucket(path, recs, frame=700, block=None):
    stream = bytearray()
    for cols, fl, bases in recs:
        stream += _enc_varint(len(cols))
        prev = 0
        for c in cols:
            stream += _enc_varint(c - prev)
            prev = c
        stream += _enc_varint((len(bases) << 4) | fl)
        packed = bytearray((len(bases) + 3) // 4)
        for i, ch in enumerate(bases):
            packed[i // 4] |= F.ACGT.index(ch) << (2 * (i % 4))
        stream += packed
    with open(path, "wb") as f:
        for o in range(0, len(stream), frame):
            blk = bytes(stream[o:o + frame])
            lit = (block or _literal_block)(blk)
            f.write(struct.pack("<II", len(blk), len(lit)) + lit)
        f.write(struct.pack("<I", 0))


@pytest.mark.parametrize("k,m,B,frames", [(31, 19, 16, "literal"), (21, 11, 8, "literal"), (31, 19, 4, "liblz4")])
def test_multicolour_records(gpu_ctx, tmp_path, k, m, B, frames):
    """Records merged across colours (sorted unique colour lists), shuffled
    within each bucket, split over literal-only LZ4 frames of 700 bytes, or
    over frames of up to 200 KB compressed by the system liblz4."""
    if frames == "liblz4" and SYS is None:
        pytest.skip("system liblz4 not present")
    files = _fasta_job(tmp_path, colors=7, glen=8000, seed=12)
    data, off = G.load_fasta(files)
    sup = O.Build(data, off, k, m, B).supers(1)
    rng = np.random.default_rng(3)
    d = tmp_path / "merged"
    d.mkdir()
    for b in range(B):
        groups = {}
        for i in np.nonzero(sup["bucket"] == b)[0]:
            s, ln = int(sup["start"][i]), int(sup["len"][i])
            key = (bytes(data[s:s + ln]).upper(), int(sup["flags"][i]))
            groups.setdefault(key, set()).add(int(sup["color"][i]))
        recs = [(tuple(sorted(cols)), fl, bases) for (bases, fl), cols in groups.items()]
        rng.shuffle(recs)
        if frames == "liblz4":
            _write_bucket(str(d / f"bucket_{b}.bin"), recs, frame=200000, block=_sys_block)
        else:
            _write_bucket(str(d / f"bucket_{b}.bin"), recs)
    _compare_phase2(gpu_ctx, tmp_path, str(d), B, k, len(files))


def test_failed_profiled_run_then_file_write(gpu_ctx, tmp_path):
    """A profiled run that fails (an injected device out-of-memory inside
    phase 1 or 2) leaves no profiler behind on its thread: the job is
    destroyed, and the next job's bucket-file write on the same thread (whose
    launches are profiled only when its own job enables it) runs normally."""
    data, off = G.synth(seed=2, num_colors=2, genome_len=5000)
    for n in (3, 25):
        with gpu_ctx.job(data, off, 31, 19, 16) as j:
            j.set_profiling(True)
            G.lib().cdbg_gpu_inject_oom(n)
            try:
                with pytest.raises(RuntimeError, match="injected"):
                    j.run()
            finally:
                G.lib().cdbg_gpu_inject_oom(-1)
    d = str(tmp_path / "b")
    with gpu_ctx.job(data, off, 31, 19, 16) as j2:
        j2.run()
        j2.write_bucket_files(d)
    ob = O.Build(data, off, 31, 19, 16)
    sup = ob.supers(1)
    boff = np.concatenate([[0], np.cumsum(np.bincount(sup["bucket"], minlength=16))])
    for b in range(16):
        assert _records(os.path.join(d, f"bucket_{b}.bin")) == _merged_bucket(data, sup, boff[b], boff[b + 1])


def _stream(recs, varint=_enc_varint):
    out = bytearray()
    for cols, fl, bases in recs:
        out += varint(len(cols))
        prev = 0
        for c in cols:
            out += varint(c - prev)
            prev = c
        out += varint((len(bases) << 4) | fl)
        packed = bytearray((len(bases) + 3) // 4)
        for i, ch in enumerate(bases):
            packed[i // 4] |= F.ACGT.index(ch) << (2 * (i % 4))
        out += packed
    return bytes(out)


def _write_stream(path, stream, frame=700):
    with open(path, "wb") as f:
        for o in range(0, len(stream), frame):
            blk = stream[o:o + frame]
            lit = _literal_block(blk)
            f.write(struct.pack("<II", len(blk), len(lit)) + lit)
        f.write(struct.pack("<I", 0))


def _padded_varint(width):
    """LEB128 with redundant continuation bytes, `width` bytes long (at most
    ten: varint_read takes ten bytes, super_kmer.hpp:71-82)."""
    def enc(v):
        out = bytearray()
        for i in range(width):
            b = v & 0x7F
            v >>= 7
            out.append(b | (0x80 if i + 1 < width else 0))
        return bytes(out)
    return enc


def test_stream_edge_cases(gpu_ctx, tmp_path):
    """Streams the reader must walk as read_super_kmer does
    (super_kmer.hpp:121-143): varints padded to 2..10 bytes, a record of
    many colours, records shorter than k, a stream cut inside its last record
    (reading stops there), an empty bucket, a bucket without end marker."""
    k, m, B, nc = 31, 19, 6, 5
    files = _fasta_job(tmp_path, colors=nc, glen=6000, seed=31)
    data, off = G.load_fasta(files)
    sup = O.Build(data, off, k, m, B).supers(1)
    boff = np.concatenate([[0], np.cumsum(np.bincount(sup["bucket"], minlength=B))])
    d = tmp_path / "edge"
    d.mkdir()
    for b in range(B):
        recs = _merged_bucket(data, sup, boff[b], boff[b + 1])
        path = str(d / f"bucket_{b}.bin")
        if b == 0:
            _write_stream(path, _stream(recs, _padded_varint(10)))
        elif b == 1:
            s = b"".join(_stream([r], _padded_varint(2 + i % 8)) for i, r in enumerate(recs))
            _write_stream(path, s, frame=333)
        elif b == 2:
            s = _stream(recs + [((0, 1, 2, 3, 4), 3, b"ACGTACGTAC"), ((1,), 0, b"A" * 31)])
            _write_stream(path, s[:-3])  # the last record loses its last bytes
        elif b == 3:
            _write_stream(path, b"")
        elif b == 4:
            s = _stream([((0,) * 200 + (1, 2, 3, 4), 12, b"C" * 5)] + recs)  # 204 colours, len < k
            _write_stream(path, s)
            with open(path, "r+b") as f:  # drop the end marker
                f.truncate(os.path.getsize(path) - 4)
        else:
            _write_stream(path, _stream(recs), frame=64)
    _compare_phase2(gpu_ctx, tmp_path, str(d), B, k, nc)


def test_colour_out_of_range(gpu_ctx, tmp_path):
    d = tmp_path / "bad"
    d.mkdir()
    _write_stream(str(d / "bucket_0.bin"), _stream([((0, 2), 0, b"ACGT" * 10)]))
    _write_stream(str(d / "bucket_1.bin"), _stream([((1, 7), 0, b"ACGA" * 10)]))
    with pytest.raises(RuntimeError, match="bucket 1 names colour 7"):
        gpu_ctx.job_from_bucket_files(str(d), 2, 31, 4)

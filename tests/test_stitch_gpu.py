"""Phase 3 (stitch) on the device, and the parallel record parser it shares
with the host stitch.

* cdbg_frag_offsets: the record starts of a frag_unitigs.bin stream
  (frag_record_parse, unitig_spill.hpp:646-665, record after record) from a
  parse split over threads that guess where their range's first record
  starts -- equal to a plain sequential parse for every thread count, on
  streams whose colour-run bytes are built to look like record starts, and
  the same error on a truncated stream (CPU).
* cdbg_gpu_stitch: the same unitigs, in the same order, with the same
  statistics as the host stitch (whose own tests pin it to the ground-truth
  coloured compacted dBG) -- on the oracle's fragments of mutated genomes,
  repeats, cycles, hairpins and branches, k = 11 ... 63 (64- and 128-bit
  keys), skewed bucket maps, fragments longer than one copy piece, on the
  fragment stream of a device build of a synthetic pangenome, on a file; the
  same errors on malformed streams (GPU).
"""
import struct

import numpy as np
import pytest

import oracle as O
from conftest import cdbg_gpu as G
from test_oracle_phase2 import _job, _mutated

FOREIGN = (1 << 64) - 1


def _rec(open_, runs, seq):
    b = struct.pack("<BI", open_, len(runs))
    for cid, n in runs:
        b += struct.pack("<QI", cid, n)
    return b + struct.pack("<I", len(seq)) + seq


def _seq_parse(buf):
    off, p, n = [], 0, len(buf)
    while p < n:
        if n - p < 9:
            raise ValueError("truncated")
        nr = struct.unpack_from("<I", buf, p + 1)[0]
        if n - p < 9 + 12 * nr:
            raise ValueError("truncated")
        ln = struct.unpack_from("<I", buf, p + 5 + 12 * nr)[0]
        if n - p - 9 - 12 * nr < ln:
            raise ValueError("truncated")
        off.append(p)
        p += 9 + 12 * nr + ln
    return off


def _synthetic_stream(seed, n_recs, max_len=300, max_runs=6, long_every=0):
    """Random records whose colour ids hold bytes that look like record
    starts (an ACGT byte, then 0-3, then small lengths)."""
    r = np.random.default_rng(seed)
    out = bytearray()
    for i in range(n_recs):
        nr = int(r.integers(0, max_runs + 1))
        runs = []
        for _ in range(nr):
            if r.random() < 0.5:
                cid = int.from_bytes(bytes([b"ACGT"[int(r.integers(0, 4))], int(r.integers(0, 4)),
                                            int(r.integers(0, 3)), 0, 0, 0, 0, 0]), "little")
            else:
                cid = int(r.integers(0, 1 << 62))
            runs.append((cid, int(r.integers(0, 50))))
        ln = int(r.integers(0, max_len))
        if long_every and i % long_every == 0:
            ln = 200_000
        seq = bytes(np.frombuffer(b"ACGT", np.uint8)[r.integers(0, 4, ln)])
        out += _rec(int(r.integers(0, 4)), runs, seq)
    return bytes(out)


@pytest.mark.parametrize("seed,n_recs,long_every", [(1, 20000, 0), (2, 15000, 97), (3, 3000, 5)])
def test_frag_offsets_parallel_equals_sequential(seed, n_recs, long_every):
    buf = _synthetic_stream(seed, n_recs, long_every=long_every)
    want = _seq_parse(buf)
    for th in (1, 2, 3, 7, 16, 64):
        got = G.frag_offsets(buf, threads=th)
        assert got.tolist() == want, th


def test_frag_offsets_truncated_stream_fails_for_every_thread_count():
    buf = _synthetic_stream(4, 20000)
    for cut in (1, 5, 20, len(buf) // 2 + 3):
        bad = buf[:-cut] if cut < len(buf) // 2 else buf[:cut]
        try:
            _seq_parse(bad)
            continue  # the cut happens to fall on a record boundary
        except ValueError:
            pass
        for th in (1, 4, 16):
            with pytest.raises(RuntimeError, match="truncated"):
                G.frag_offsets(bad, threads=th)


# --------------------------------------------------------------------------- GPU

@pytest.fixture(scope="module")
def ctx():
    c = G.Context(0)
    yield c
    c.close()


def _same(ctx, frags, k, threads=4):
    with G.stitch(frags, k, threads=threads) as H:
        want, wst = H.unitigs(), H.stats()
    with ctx.stitch(frags, k, threads=threads) as D:
        got, gst = D.unitigs(), D.stats()
    assert gst == wst
    assert got == want
    return gst


def _oracle_frags(data, off, k, m, B, cuts=None):
    return O.Build(data, off, k, m, B, cuts=cuts).frags()


@pytest.mark.gpu
@pytest.mark.parametrize("k,m,B,seed", [(5, 2, 2, 9), (7, 3, 4, 8), (15, 7, 4, 1), (21, 11, 8, 3), (31, 19, 64, 4),
                                        (32, 19, 16, 5), (33, 21, 8, 6), (63, 27, 16, 7)])
def test_gpu_stitch_equals_host_mutated_genomes(ctx, k, m, B, seed):
    """k = 5 and 7 make dense graphs (every canonical 5-mer occurs); k = 31 is
    the widest 64-bit key, 32 the narrowest 128-bit one."""
    base = bytes(np.frombuffer(b"ACGT", np.uint8)[np.random.default_rng(seed).integers(0, 4, 3000)])
    data, off = _job([[g] for g in _mutated(base, 4, 0.02, 0.003, seed + 100)])
    st = _same(ctx, _oracle_frags(data, off, k, m, B), k)
    assert st["joins"] > 0 and st["num_unitigs"] > 10


@pytest.mark.gpu
@pytest.mark.parametrize("k,m,B", [(11, 5, 4), (15, 7, 8), (12, 5, 3), (11, 5, 1)])
def test_gpu_stitch_equals_host_repeats_cycles_hairpins(ctx, k, m, B):
    r = np.random.default_rng(5)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    unit = bytes(acgt[r.integers(0, 4, 40)])
    rc = lambda s: s[::-1].translate(bytes.maketrans(b"ACGT", b"TGCA"))
    shared = bytes(acgt[r.integers(0, 4, 300)])
    ring = bytes(acgt[r.integers(0, 4, 200)])
    ring2 = bytes(acgt[r.integers(0, 4, 90)])
    colours = [
        [unit * 6, b"acgtNacgtacgtacgtacgtacgt" + shared, ring + ring[:k - 1]],
        [unit + rc(unit), shared[:150] + b"A" + shared[150:] + b"NN" + shared[:150] + b"C" + shared[150:],
         ring[:120]],
        [unit * 2 + shared, b"GGGGGGGGGGGGGGGGGGGGGG", ring2 + ring2[:k - 1]],
    ]
    data, off = _job(colours)
    st = _same(ctx, _oracle_frags(data, off, k, m, B), k)
    assert st["num_unitigs"] > 0


@pytest.mark.gpu
def test_gpu_stitch_equals_host_synthetic_pangenome_and_skew(ctx):
    data, off = G.synth(seed=9, num_colors=5, genome_len=4000, snp_rate=0.01, indel_rate=0.002,
                        pav_frac=0.05, pav_min=100, pav_max=400, n_inversions=1, inv_min=300, inv_max=800)
    _same(ctx, _oracle_frags(data, off, 21, 9, 16), 21)
    _same(ctx, _oracle_frags(data, off, 31, 19, 256, cuts=G.zipf_bucket_cuts(256, 1.5)), 31)
    _same(ctx, _oracle_frags(data, off, 45, 19, 64), 45)


@pytest.mark.gpu
def test_gpu_stitch_long_fragments(ctx):
    """Two identical genomes: one colour set, unitigs of 150 kbp, longer than
    a 4096-byte copy piece and than a 64 KiB parse chunk; one bucket (whole
    unitigs in one fragment) and four (chains of fragments)."""
    data, off = G.synth(seed=3, num_colors=2, genome_len=150000, snp_rate=0.0, indel_rate=0.0)
    for k, m, B in [(31, 15, 1), (31, 15, 4), (63, 25, 4)]:
        f = _oracle_frags(data, off, k, m, B)
        _same(ctx, f, k)
        with G.stitch(f, k) as H:
            assert max(len(s) for s, _ in H.unitigs()) > 3 * 4096


@pytest.mark.gpu
def test_gpu_stitch_device_build_pangenome(ctx, tmp_path):
    """The fragment stream of a device build (phases 1-2) of a 40-colour
    pangenome, through the host and the device stitch, and through the file
    entry point."""
    data, off = G.synth(seed=21, num_colors=40, genome_len=200_000, snp_rate=0.01, indel_rate=0.001)
    with ctx.job(data, off, 31, 19, 256) as j:
        j.run()
        frags = j.frags()
    st = _same(ctx, frags, 31, threads=8)
    assert st["num_fragments"] > 10000 and st["joins"] > 0
    p = tmp_path / "frag_unitigs.bin"
    np.asarray(frags, np.uint8).tofile(p)
    with G.stitch_file(str(p), 31) as H, ctx.stitch_file(str(p), 31, threads=8) as D:
        assert D.stats() == H.stats()
        assert D.export()[0] == H.export()[0]


@pytest.mark.gpu
@pytest.mark.parametrize("spill", [False, True])
def test_gpu_job_stitch(ctx, monkeypatch, tmp_path, spill):
    """The stitch of a resident job's fragment stream where it lies (HBM, or
    host memory when phase 2 keeps it there) equals the host stitch of the
    same stream; the job can still write its colour-set files without the
    fragment file."""
    if spill:
        monkeypatch.setenv("CDBG_FRAG_SPILL", "1")
    data, off = G.synth(seed=23, num_colors=12, genome_len=120_000, snp_rate=0.01, indel_rate=0.001)
    with ctx.job(data, off, 31, 19, 64) as j:
        j.run()
        frags = j.frags()
        with j.stitch(threads=8) as D:
            got, gst = D.unitigs(), D.stats()
        state = j.write_phase2_files(None, str(tmp_path / "o.color_sets"))
    with G.stitch(frags, 31, threads=8) as H:
        assert H.stats() == gst
        assert H.unitigs() == got
    assert gst["num_fragments"] == state["num_fragments"] > 1000
    assert (tmp_path / "o.color_sets").exists() and not (tmp_path / "frag_unitigs.bin").exists()


@pytest.mark.gpu
def test_gpu_stitch_record_starts_it_cannot_guess(ctx):
    """Closed fragments whose sequences start with N: no record start looks
    like one (an open-flag byte before a base), so every parse chunk is
    walked again from where the true chain enters it."""
    r = np.random.default_rng(11)
    k, recs = 11, []
    for i in range(6000):
        ln = int(r.integers(k, 600))
        seq = b"N" + bytes(np.frombuffer(b"ACGT", np.uint8)[r.integers(0, 4, ln - 1)])
        recs.append(_rec(0, [(i % 7, ln - k + 1)], seq))
    buf = b"".join(recs)
    assert len(buf) > 20 * 65536
    st = _same(ctx, buf, k)
    assert st["num_fragments"] == 6000 and st["num_unitigs"] == 6000 and st["joins"] == 0


@pytest.mark.gpu
def test_gpu_stitch_out_of_memory_is_an_error_and_leaves_nothing_behind(ctx):
    """A device allocation that fails inside the stitch (injected) is reported
    as an error; the next stitch on the same context gives the full result."""
    data, off = G.synth(seed=9, num_colors=5, genome_len=4000, snp_rate=0.01, indel_rate=0.002)
    f = _oracle_frags(data, off, 21, 9, 16)
    for n in (0, 3, 12):
        G.lib().cdbg_gpu_trim_cache()
        G.lib().cdbg_gpu_inject_oom(n)
        try:
            with pytest.raises(RuntimeError, match="out of memory"):
                with ctx.stitch(f, 21):
                    pass
        finally:
            G.lib().cdbg_gpu_inject_oom(-1)
    _same(ctx, f, 21)


@pytest.mark.gpu
def test_gpu_stitch_empty_stream(ctx):
    with ctx.stitch(b"", 31) as D:
        st = D.stats()
        assert st["num_fragments"] == 0 and st["num_unitigs"] == 0 and D.unitigs() == []


def _tig_pieces(k, seed):
    """A sequence with distinct k-mers and three fragments that tile it: f0
    open right, f1 open both sides, f2 open left (forward orientation)."""
    r = np.random.default_rng(seed)
    while True:
        t = bytes(np.frombuffer(b"ACGT", np.uint8)[r.integers(0, 4, 40)])
        kms = [t[i:i + k] for i in range(len(t) - k + 1)]
        rc = lambda s: s[::-1].translate(bytes.maketrans(b"ACGT", b"TGCA"))
        if len(set(kms) | set(rc(x) for x in kms)) == 2 * len(kms):
            break
    a, b = 10, 24
    return t, (t[:a + k], t[a:b + k], t[b:])


@pytest.mark.gpu
def test_gpu_stitch_same_errors_as_host(ctx):
    k = 7
    t, (s0, s1, s2) = _tig_pieces(k, 1)
    n = lambda s: len(s) - k + 1
    good = _rec(2, [(1, n(s0))], s0) + _rec(3, [(2, n(s1))], s1) + _rec(1, [(1, n(s2))], s2)
    st = _same(ctx, good, k)
    assert st["joins"] == 2 and st["num_unitigs"] == 3
    # a foreign middle fragment: its k-mers stay inside the tig
    inner = _rec(2, [(1, n(s0))], s0) + _rec(3, [(FOREIGN, n(s1))], s1) + _rec(1, [(1, n(s2))], s2)
    # foreign at both ends only: dropped
    ends = _rec(2, [(FOREIGN, n(s0))], s0) + _rec(3, [(5, n(s1))], s1) + _rec(1, [(FOREIGN, n(s2))], s2)
    st = _same(ctx, ends, k)
    assert st["foreign_dropped"] > 0
    cases = {
        "foreign k-mer inside": inner,
        "do not cover": _rec(2, [(1, n(s0) + 1)], s0) + _rec(1, [(1, n(s1))], s1),
        "shorter than k": _rec(0, [(1, 1)], s0[:k - 1]),
        "non-ACGT": _rec(2, [(1, n(s0))], s0[:-1] + b"N") + _rec(1, [(1, n(s1))], s1),
        "more than two": _rec(2, [(1, n(s0))], s0) + _rec(1, [(1, n(s1))], s1) + _rec(1, [(3, n(s1))], s1),
        "truncated": good[:-2],
    }
    for msg, buf in cases.items():
        with pytest.raises(RuntimeError, match=msg):
            with G.stitch(buf, k):
                pass
        with pytest.raises(RuntimeError, match=msg):
            with ctx.stitch(buf, k):
                pass

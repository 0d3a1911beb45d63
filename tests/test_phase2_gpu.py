"""Phase 2 on the GPU, bit-exact against the oracle (the reference's
process_buckets semantics at -t 1 over the same record stream):

* the frag_unitigs.bin byte stream (every fragment's open flags, colour runs
  with global cids, sequence, in (bucket, seed rank) order);
* the colour-set bit stream and per-class bit offsets;
* num_kmers, node, fragment and colour-set counts.
"""
import numpy as np
import pytest

import oracle as O
from conftest import cdbg_gpu as G

pytestmark = pytest.mark.gpu
ACGT = np.frombuffer(b"ACGT", np.uint8)


def _run_both(ctx, data, off, k, m, B, chunk=0, cuts=None):
    ob = O.Build(data, off, k, m, B, threads=4, cuts=cuts)
    with ctx.job(data, off, k, m, B) as job:
        if cuts is not None:
            job.set_bucket_map(cuts)
        if chunk:
            job.set_chunk_tuples(chunk)
        job.run()
        if chunk:
            assert job.num_chunks() > 1
        s = job.stats()
        gf = job.frags()
        gw, go = job.color_sets()
    return ob, s, gf, gw, go


def _assert_same(ob, s, gf, gw, go):
    assert s["num_kmers"] == ob.stat("num_kmers")
    assert s["num_nodes"] == ob.stat("num_nodes")
    assert s["num_frags"] == ob.stat("num_frags")
    assert s["num_color_sets"] == ob.stat("num_sets")
    assert s["cs_bits"] == ob.stat("cs_bits")
    assert s["total_integers"] == ob.stat("total_integers")
    of = ob.frags()
    if not np.array_equal(gf, of):
        n = min(len(gf), len(of))
        first = int(np.argmax(gf[:n] != of[:n])) if np.any(gf[:n] != of[:n]) else n
        raise AssertionError(f"frag stream differs (gpu {len(gf)} B, oracle {len(of)} B), first byte {first}")
    ow, oo = ob.color_sets()
    assert np.array_equal(go, oo), "colour-set offsets differ"
    assert np.array_equal(gw, ow), "colour-set words differ"


@pytest.mark.parametrize("k,m,B,colors,glen", [
    (31, 19, 64, 10, 20000),     # the small plain configuration, shrunk
    (31, 21, 1024, 6, 30000),
    (31, 12, 16, 3, 15000),
    (63, 27, 256, 5, 20000),     # 128-bit k-mers
    (32, 19, 128, 5, 20000),     # 128-bit k-mers whose high word is empty
    (33, 21, 96, 5, 20000),      # one base in the high word
    (47, 23, 200, 5, 20000),
    (15, 9, 37, 4, 8000),
    (21, 11, 1, 2, 5000),        # one bucket: no cross-bucket boundaries
])
def test_phase2_matches_oracle(gpu_ctx, k, m, B, colors, glen):
    data, off = G.synth(seed=3, num_colors=colors, genome_len=glen, snp_rate=0.01, indel_rate=0.001)
    _assert_same(*_run_both(gpu_ctx, data, off, k, m, B))


@pytest.mark.parametrize("num_colors", [1, 3, 8, 40])
def test_phase2_codec_branches(gpu_ctx, num_colors):
    """Colour counts that drive the sparse, bitmap and complement branches."""
    data, off = G.synth(seed=11, num_colors=num_colors, genome_len=6000, snp_rate=0.02, indel_rate=0.002)
    _assert_same(*_run_both(gpu_ctx, data, off, 31, 19, 32))


def test_phase2_repeats_and_cycles(gpu_ctx):
    """Tandem repeats, homopolymers and reverse-complement palindromes make
    branching nodes, cycles and contig breaks inside paths."""
    rng = np.random.default_rng(9)
    unit = ACGT[rng.integers(0, 4, 37)].tobytes()
    pal = ACGT[rng.integers(0, 4, 40)].tobytes()
    rc = pal[::-1].translate(bytes.maketrans(b"ACGT", b"TGCA"))
    parts = [
        unit * 30 + ACGT[rng.integers(0, 4, 500)].tobytes(),
        b"A" * 300 + b"C" * 50 + b"AC" * 100,
        pal + rc + pal + rc,
        ACGT[rng.integers(0, 4, 2000)].tobytes() + unit * 5,
    ]
    data = b"".join(p + b"\n" for p in parts)
    off = np.zeros(len(parts) + 1, np.uint64)
    off[1:] = np.cumsum([len(p) + 1 for p in parts])
    data = np.frombuffer(data, np.uint8).copy()
    for (k, m, B) in [(31, 19, 8), (15, 7, 3), (9, 5, 1), (63, 27, 5)]:
        _assert_same(*_run_both(gpu_ctx, data, off, k, m, B))


def test_phase2_rerun_same_job(gpu_ctx):
    data, off = G.synth(seed=4, num_colors=4, genome_len=10000)
    ob = O.Build(data, off, 31, 19, 64)
    with gpu_ctx.job(data, off, 31, 19, 64) as job:
        for _ in range(2):
            job.run()
            s = job.stats()
            _assert_same(ob, s, job.frags(), *job.color_sets())


def test_phase2_distinct_kmers_bruteforce(gpu_ctx):
    """num_kmers equals the number of distinct canonical k-mers (verify.py:103-131)."""
    data, off = G.synth(seed=5, num_colors=3, genome_len=3000)
    k = 21
    kms = set()
    for c in range(3):
        seq = bytes(data[off[c]:off[c + 1]])
        for run in seq.replace(b"\n", b"N").split(b"N"):
            for i in range(len(run) - k + 1):
                s = run[i:i + k]
                r = s[::-1].translate(bytes.maketrans(b"ACGT", b"TGCA"))
                kms.add(min(s, r))
    with gpu_ctx.job(data, off, k, 11, 16) as job:
        job.run()
        assert job.stats()["num_kmers"] == len(kms)


@pytest.mark.parametrize("k,m,B,chunk", [(31, 19, 64, 20000), (31, 19, 64, 1), (63, 27, 97, 30000),
                                         (31, 21, 1024, 50000), (32, 19, 80, 30000)])
def test_phase2_chunked(gpu_ctx, k, m, B, chunk):
    """Buckets processed in several chunks: colour-set ids, the colour-set
    stream and the fragment stream continue across chunks."""
    data, off = G.synth(seed=8, num_colors=7, genome_len=25000, snp_rate=0.015, indel_rate=0.001)
    _assert_same(*_run_both(gpu_ctx, data, off, k, m, B, chunk=chunk))


@pytest.mark.parametrize("k,m,B", [(31, 19, 4), (63, 27, 2)])
def test_phase2_large_buckets_split_into_cells(gpu_ctx, k, m, B):
    """Buckets of ~10^5 occurrences: the tuple sort splits them on their top
    key bits into cells that are finished in LDS."""
    data, off = G.synth(seed=14, num_colors=8, genome_len=120000, snp_rate=0.01, indel_rate=0.001)
    _assert_same(*_run_both(gpu_ctx, data, off, k, m, B))


@pytest.mark.parametrize("k,m,cell_target,persistent", [
    (31, 19, 300, 0),   # 11 bits in the global passes: every class packs (51 key bits + 13 position bits)
    (31, 19, 64, 1),    # 13 global bits: 49 key bits
    (31, 19, 1200, 0),  # 9 global bits: 53 key bits + 11 position bits fill the 2048-key class's word
    (31, 19, 2400, 1),  # 8 global bits: 54 + 10 fill the 1024-key class's word; larger cells rank
    (63, 27, 300, 0),   # 126 - 11 = 115 key bits + 13 position bits: a full 128-bit word
    (63, 27, 64, 1),
    (63, 27, 2400, 0),  # 118 key bits: the 512- and 1024-key classes pack (9 and 10 position bits)
])
def test_phase2_packed_cells_at_widest_key_bits(gpu_ctx, k, m, cell_target, persistent):
    """One bucket of ~3.6e5 occurrences with small cells: a cell's unsorted
    key bits and its position (log2 of the class's capacity in bits) fit one
    key word, so the cells are merge-sorted as packed words, up to words
    filled to the last bit.  Cells whose bits do not fit take the ranking
    passes."""
    data, off = G.synth(seed=21, num_colors=6, genome_len=60000, snp_rate=0.01, indel_rate=0.001)
    ob = O.Build(data, off, k, m, 1, threads=4)
    with gpu_ctx.job(data, off, k, m, 1) as job:
        job.set_tuning(cell_target=cell_target, persistent=persistent)
        job.run()
        s = job.stats()
        gf = job.frags()
        gw, go = job.color_sets()
    d = {300: 11, 1200: 9, 2400: 8}.get(cell_target)  # the global bits these sizes give
    assert d is None or 2 ** (d - 1) * cell_target < s["num_occ"] <= 2 ** d * cell_target
    _assert_same(ob, s, gf, gw, go)


@pytest.mark.parametrize("k,m,B", [(31, 19, 8), (63, 27, 4), (21, 11, 16)])
def test_phase2_heavy_kmers(gpu_ctx, k, m, B):
    """One k-mer repeated far beyond what a cell holds in LDS (homopolymer
    and short tandem repeats in every colour): those cells take the global
    sort."""
    rng = np.random.default_rng(8)
    seqs = []
    for c in range(3):
        body = ACGT[rng.integers(0, 4, 20000)].tobytes()
        seqs.append(b"A" * 12000 + body[:10000] + b"ACG" * 4000 + body[10000:] + b"\n")
    data = np.frombuffer(b"".join(seqs), np.uint8).copy()
    off = np.concatenate([[0], np.cumsum([len(x) for x in seqs])]).astype(np.uint64)
    _assert_same(*_run_both(gpu_ctx, data, off, k, m, B))


@pytest.mark.parametrize("k,m,B", [(31, 21, 1024), (63, 27, 16)])
def test_phase2_fragment_sort_64bit_keys(gpu_ctx, k, m, B, monkeypatch):
    """Fragments are ordered with 32-bit (local bucket, seed rank) keys when
    they fit; the 64-bit key path (large buckets) orders them the same."""
    monkeypatch.setenv("CDBG_FRAG_KEY64", "1")
    data, off = G.synth(seed=5, num_colors=5, genome_len=20000, snp_rate=0.01, indel_rate=0.001)
    _assert_same(*_run_both(gpu_ctx, data, off, k, m, B))


@pytest.mark.parametrize("stage", ["0", "1"])
@pytest.mark.parametrize("k,m,B", [(31, 21, 256), (63, 27, 64)])
def test_phase2_fragment_writer_paths(gpu_ctx, k, m, B, stage, monkeypatch):
    """Fragment records are built from positions staged in LDS when a block's
    positions fit the stage (many short fragments: 60 colours at 3% SNPs),
    else from HBM; forced off, every block reads HBM.  Both give the
    oracle's bytes."""
    if stage == "0":
        monkeypatch.setenv("CDBG_FRAG_NOSTAGE", "1")
    data, off = G.synth(seed=21, num_colors=60, genome_len=4000, snp_rate=0.03, indel_rate=0.002)
    _assert_same(*_run_both(gpu_ctx, data, off, k, m, B))


@pytest.mark.parametrize("k,m,B,s,chunk", [(31, 19, 512, 1.1, 0), (63, 27, 256, 1.1, 0), (31, 21, 1024, 1.3, 40000),
                                           (32, 19, 300, 1.1, 0)])
def test_phase2_skewed_bucket_map(gpu_ctx, k, m, B, s, chunk):
    """Zipf bucket sizes (a bucket map): one bucket holds a large share of
    the occurrences, most hold a few; results equal the oracle under the same
    map, also in several chunks."""
    data, off = G.synth(seed=12, num_colors=6, genome_len=30000, snp_rate=0.01, indel_rate=0.001)
    _assert_same(*_run_both(gpu_ctx, data, off, k, m, B, chunk=chunk, cuts=G.zipf_bucket_cuts(B, s)))


@pytest.mark.parametrize("k,m,B,chunk", [(31, 19, 64, 20000), (63, 27, 97, 30000)])
def test_phase2_fragment_stream_in_host_memory(gpu_ctx, k, m, B, chunk, monkeypatch):
    """A job whose fragment stream would not fit HBM keeps it in host memory
    chunk by chunk (forced here after the first chunk): the same bytes, from
    the job and from its files."""
    monkeypatch.setenv("CDBG_FRAG_SPILL", "1")
    data, off = G.synth(seed=8, num_colors=7, genome_len=25000, snp_rate=0.015, indel_rate=0.001)
    _assert_same(*_run_both(gpu_ctx, data, off, k, m, B, chunk=chunk))

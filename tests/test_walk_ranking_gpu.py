"""Long paths ranked by pointer jumping (the walk of bucket_walker.hpp:451-566
for paths longer than the end walks' step cap), bit-exact against the oracle.

With one or two buckets a bucket's local graph keeps whole unitigs, so a
genome of long unique stretches and a long tandem repeat gives paths of
tens of thousands of nodes; CDBG_WALK_CAP (a test switch) lowers the cap so
that ordinary workloads, hairpins and cycles go through the ranking too.
"""
import numpy as np
import pytest

import oracle as O
from test_phase2_gpu import ACGT, _assert_same, _run_both
from conftest import cdbg_gpu as G

pytestmark = pytest.mark.gpu


def _tandem_genome(seed, unique=40000, unit=2500, copies=12, snp_every=0):
    rng = np.random.default_rng(seed)
    a = ACGT[rng.integers(0, 4, unique)].tobytes()
    u = ACGT[rng.integers(0, 4, unit)].tobytes()
    b = ACGT[rng.integers(0, 4, unique)].tobytes()
    g = bytearray(a + u * copies + b)
    variants = [bytes(g)]
    if snp_every:
        v = bytearray(g)
        for i in range(snp_every // 2, len(v), snp_every):
            v[i] = int(ACGT[(int(np.searchsorted(ACGT, v[i])) + 1) % 4])
        variants.append(bytes(v))
    parts = [x + b"\n" for x in variants]
    data = np.frombuffer(b"".join(parts), np.uint8).copy()
    off = np.concatenate([[0], np.cumsum([len(p) for p in parts])]).astype(np.uint64)
    return data, off


def _ranked(ctx, data, off, k, m, B):
    """The job's kernel names when run with profiling on."""
    with ctx.job(data, off, k, m, B) as job:
        job.set_profiling(True)
        job.run()
        return set(job.kernel_times())


@pytest.mark.parametrize("k,m,B", [(31, 19, 1), (31, 19, 2), (63, 27, 1)])
def test_long_tandem_repeat_paths(gpu_ctx, k, m, B):
    """Unique stretches of 40 kbp around 12 copies of a 2.5 kbp unit, two
    colours (one with a SNP every 5 kbp): with one bucket, paths of up to
    ~40,000 nodes (ranked); with two, consecutive super-k-mers share a bucket
    half the time and paths stay below the cap (walked)."""
    data, off = _tandem_genome(5, snp_every=5000)
    _assert_same(*_run_both(gpu_ctx, data, off, k, m, B))
    names = _ranked(gpu_ctx, data, off, k, m, B)
    assert (("k_long_jump" in names) and ("k_walk_frags_long" in names)) == (B == 1)


def test_long_paths_chunked(gpu_ctx, monkeypatch):
    """One colour, two buckets in two chunks, ends stopped after 64 links:
    ranked and walked paths in each chunk."""
    monkeypatch.setenv("CDBG_WALK_CAP", "64")
    data, off = _tandem_genome(6, unique=30000, unit=1800, copies=20)
    _assert_same(*_run_both(gpu_ctx, data, off, 31, 21, 2, chunk=20000))


@pytest.mark.parametrize("cap", ["1", "3", "9"])
@pytest.mark.parametrize("k,m,B,colors,glen", [
    (31, 19, 64, 10, 20000),
    (63, 27, 256, 5, 20000),
    (32, 19, 128, 5, 20000),
    (21, 11, 1, 2, 5000),
])
def test_forced_ranking_matches_oracle(gpu_ctx, monkeypatch, cap, k, m, B, colors, glen):
    """Every path longer than `cap` links is ranked: short paths, contig
    breaks inside ranked paths (several seeds per path) and chunking."""
    monkeypatch.setenv("CDBG_WALK_CAP", cap)
    data, off = G.synth(seed=3, num_colors=colors, genome_len=glen, snp_rate=0.01, indel_rate=0.001)
    _assert_same(*_run_both(gpu_ctx, data, off, k, m, B))


@pytest.mark.parametrize("cap", ["1", "4"])
def test_forced_ranking_repeats_cycles_hairpins(gpu_ctx, monkeypatch, cap):
    """Tandem repeats, homopolymers and reverse-complement palindromes
    (hairpin paths go to the general loop, cycles to the cycle loop)."""
    monkeypatch.setenv("CDBG_WALK_CAP", cap)
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


def test_forced_ranking_chunked_skewed(gpu_ctx, monkeypatch):
    monkeypatch.setenv("CDBG_WALK_CAP", "2")
    data, off = G.synth(seed=12, num_colors=6, genome_len=30000, snp_rate=0.01, indel_rate=0.001)
    _assert_same(*_run_both(gpu_ctx, data, off, 31, 21, 1024, chunk=40000, cuts=G.zipf_bucket_cuts(1024, 1.3)))


def _kernels(ctx, data, off, k, m, B, cuts=None):
    with ctx.job(data, off, k, m, B) as job:
        if cuts is not None:
            job.set_bucket_map(cuts)
        job.set_profiling(True)
        job.run()
        return set(job.kernel_times())


@pytest.mark.parametrize("k,m", [(31, 19), (63, 27)])
def test_long_fragment_colour_runs(gpu_ctx, k, m):
    """One bucket; colours that cover different stretches of one genome make
    a fragment of ~60,000 k-mers with several colour runs (records of more
    than 4096 k-mers are sized and written by a block each)."""
    rng = np.random.default_rng(17)
    g = ACGT[rng.integers(0, 4, 60000)].tobytes()
    parts = [g + b"\n", g[:30000] + b"\n", g[20000:50000] + b"\n", g[:200000] + b"\n"]
    data = np.frombuffer(b"".join(parts), np.uint8).copy()
    off = np.concatenate([[0], np.cumsum([len(p) for p in parts])]).astype(np.uint64)
    _assert_same(*_run_both(gpu_ctx, data, off, k, m, 1))
    assert "k_frag_write_long" in _kernels(gpu_ctx, data, off, k, m, 1)


def test_long_fragments_foreign_runs(gpu_ctx):
    """Two buckets, one holding all but 1/4096 of the minimizers: paths and
    fragments of up to ~5,800 k-mers whose fragments end on foreign boundary
    k-mers (COLOR_RUN_FOREIGN runs)."""
    data, off = G.synth(seed=19, num_colors=3, genome_len=100000, snp_rate=0.0002, indel_rate=0.00005)
    cuts = G.zipf_bucket_cuts(2, 12.0)
    _assert_same(*_run_both(gpu_ctx, data, off, 31, 19, 2, cuts=cuts))


@pytest.mark.parametrize("k,m", [(31, 21), (63, 27)])
def test_one_fragment_whole_genome(gpu_ctx, k, m):
    """A single 200 kbp colour in one bucket: one path and one fragment of
    ~200,000 k-mers."""
    rng = np.random.default_rng(23)
    g = ACGT[rng.integers(0, 4, 200000)].tobytes() + b"\n"
    data = np.frombuffer(g, np.uint8).copy()
    off = np.array([0, len(g)], np.uint64)
    ob, s, gf, gw, go = _run_both(gpu_ctx, data, off, k, m, 1)
    assert s["num_frags"] == 1
    _assert_same(ob, s, gf, gw, go)

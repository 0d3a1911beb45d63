"""Phase 1 on the GPU, bit-exact against the oracle: every super-k-mer
(start, length, flags, bucket, colour) in position order, the bucket-major
order and the bucket offsets."""
import numpy as np
import pytest

import oracle as O
from conftest import cdbg_gpu as G

pytestmark = pytest.mark.gpu
ACGT = np.frombuffer(b"ACGT", np.uint8)


def _edge_case_job(rng):
    """Colours that exercise tile borders, non-ACGT bytes, case, empty colours."""
    parts = [
        ACGT[rng.integers(0, 4, 9000)].tobytes(),                 # > 2 tiles
        b"N" * 5000 + ACGT[rng.integers(0, 4, 50)].tobytes(),
        b"",                                                        # empty colour
        b"A" * 12000 + b"C" * 10,                                   # long homopolymer
        ACGT[rng.integers(0, 4, 3000)].tobytes().lower() + b"\r\n" + ACGT[rng.integers(0, 4, 20)].tobytes(),
        b"ACGTNACGTTGCA" * 500,
    ]
    data = b"".join(p + b"\n" for p in parts)
    off = np.zeros(len(parts) + 1, np.uint64)
    off[1:] = np.cumsum([len(p) + 1 for p in parts])
    return np.frombuffer(data, np.uint8).copy(), off


def _compare(job, ob, B):
    for which in (0, 1):
        g = job.supers(which)
        o = ob.supers(which)
        assert len(g["start"]) == len(o["start"])
        for key in ("start", "len", "flags", "bucket", "color"):
            if not np.array_equal(g[key], o[key]):
                bad = np.nonzero(g[key] != o[key])[0][:5]
                raise AssertionError(f"which={which} field={key} first mismatches at {bad}: "
                                     f"gpu={g[key][bad]} oracle={o[key][bad]}")
    o1 = ob.supers(1)
    off = job.bucket_offsets()
    counts = np.bincount(o1["bucket"], minlength=B)
    assert np.array_equal(np.diff(off), counts)


@pytest.mark.parametrize("k,m,B", [(31, 19, 64), (31, 21, 1024), (63, 27, 4096), (15, 9, 37),
                                   (5, 3, 1), (31, 12, 16384)])
def test_phase1_synthetic(gpu_ctx, k, m, B):
    data, off = G.synth(seed=1, num_colors=6, genome_len=30000, snp_rate=0.01, indel_rate=0.001)
    ob = O.Build(data, off, k, m, B, threads=4)
    with gpu_ctx.job(data, off, k, m, B) as job:
        job.phase1()
        _compare(job, ob, B)
        # a second run of the same job reuses its buffers and must agree
        job.phase1()
        _compare(job, ob, B)


@pytest.mark.parametrize("k,m,B", [(31, 19, 64), (63, 27, 97), (9, 4, 8)])
def test_phase1_edge_cases(gpu_ctx, k, m, B):
    data, off = _edge_case_job(np.random.default_rng(5))
    ob = O.Build(data, off, k, m, B)
    with gpu_ctx.job(data, off, k, m, B) as job:
        job.phase1()
        _compare(job, ob, B)


def test_phase1_no_valid_runs(gpu_ctx):
    data = np.frombuffer(b"NNNNACGTNNNN\n", np.uint8).copy()
    off = np.array([0, len(data)], np.uint64)
    with gpu_ctx.job(data, off, 31, 19, 64) as job:
        job.phase1()
        assert job.stats()["num_supers"] == 0
        assert np.all(job.bucket_offsets() == 0)


@pytest.mark.parametrize("k,m", [(20, 19), (26, 21), (27, 21), (29, 21), (30, 21), (38, 21), (39, 21),
                                 (63, 27), (63, 2)])
def test_phase1_window_widths(gpu_ctx, k, m):
    """Window widths W = k - m across every grouping of the window minimum
    (W = 1, 5: direct; 6, 8: groups of 6; 9, 17: groups of 9; 18, 36, 61:
    groups of 18) against the oracle's direct minimum."""
    data, off = G.synth(seed=7, num_colors=4, genome_len=30000, snp_rate=0.01, indel_rate=0.001)
    ob = O.Build(data, off, k, m, 64, threads=4)
    with gpu_ctx.job(data, off, k, m, 64) as job:
        job.phase1()
        _compare(job, ob, 64)


@pytest.mark.parametrize("chunk", [1, 25_000, 70_000])
def test_phase1_colour_ranges(gpu_ctx, chunk):
    """Phase 1 over ranges of whole colours (forced small ranges: one colour
    each, or a few) gives the records, bucket order and offsets of one pass,
    also with empty colours, non-ACGT bytes and a bucket map; phase 2 over
    them gives the same bytes."""
    data, off = G.synth(seed=13, num_colors=9, genome_len=20000, snp_rate=0.01, indel_rate=0.001)
    e_data, e_off = _edge_case_job(np.random.default_rng(6))
    for d, o, k, m, B, cuts in [(data, off, 31, 19, 64, None), (e_data, e_off, 31, 19, 64, None),
                                (data, off, 63, 27, 128, G.zipf_bucket_cuts(128, 1.1))]:
        ob = O.Build(d, o, k, m, B, threads=4, cuts=cuts)
        with gpu_ctx.job(d, o, k, m, B) as job:
            if cuts is not None:
                job.set_bucket_map(cuts)
            job.set_phase1_chunk_bases(chunk)
            job.run()
            if len(d) > 2 * chunk:
                assert job.num_phase1_ranges() > 1
            _compare(job, ob, B)
            assert np.array_equal(job.frags(), ob.frags())
            gw, go = job.color_sets()
            ow, oo = ob.color_sets()
            assert np.array_equal(gw, ow) and np.array_equal(go, oo)
            # the same job in one pass, then in ranges again
            job.set_phase1_chunk_bases((1 << 64) - 1)
            job.phase1()
            assert job.num_phase1_ranges() == 1
            _compare(job, ob, B)

"""Phase-2 performance knobs and the autotuner (cdbg_gpu_job_set_tuning,
cdbg_gpu_job_autotune).

No knob may change a result: every tuning below (cell sizes from 64 to 65536
keys, so cells land in every LDS class and in the global fallback; digit
widths 3 to 11; persistent cell sorts over device cell counts; their graph
capture, replayed on a second run of the job; table slack 1.25 to 16), in
one chunk and in several, gives byte-for-byte the oracle's fragment stream
and colour-set stream.  The tuner's choice is checked the same way.
"""
import numpy as np
import pytest

import oracle as O
from conftest import cdbg_gpu as G

pytestmark = pytest.mark.gpu

TUNINGS = [
    {},
    {"cell_target": 64, "digit_bits": 3},
    {"cell_target": 65536, "digit_bits": 11},
    {"persistent": 1},
    {"persistent": 1, "graph": 1},
    {"persistent": 1, "graph": 1, "cell_target": 300, "set_slack": 1.25},
    {"persistent": 1, "cell_target": 65536},
    {"set_slack": 16, "digit_bits": 1},
]


def _expected(data, off, k, m, B):
    ob = O.Build(data, off, k, m, B, threads=4)
    return ob.frags(), ob.color_sets()


def _check(job, want):
    wf, (ww, wo) = want
    f = job.frags()
    w, o = job.color_sets()
    assert np.array_equal(f, wf), "fragment stream differs"
    assert np.array_equal(o, wo) and np.array_equal(w, ww), "colour-set stream differs"


@pytest.mark.parametrize("k,m,B,chunk", [(31, 19, 64, 0), (31, 19, 64, 60000), (31, 19, 4, 0), (63, 27, 48, 0),
                                          (45, 21, 40, 50000)])
def test_knobs_keep_results(gpu_ctx, k, m, B, chunk):
    data, off = G.synth(seed=17, num_colors=8, genome_len=25000, snp_rate=0.01, indel_rate=0.001)
    want = _expected(data, off, k, m, B)
    for t in TUNINGS:
        with gpu_ctx.job(data, off, k, m, B) as job:
            job.set_tuning(**t)
            if chunk:
                job.set_chunk_tuples(chunk)
            for rep in range(2 if t.get("graph") else 1):  # the second run replays the chunks' graphs
                job.run()
                if chunk:
                    assert job.num_chunks() > 1
                _check(job, want)


def test_autotune_choice_keeps_results(gpu_ctx):
    data, off = G.synth(seed=3, num_colors=10, genome_len=40000, snp_rate=0.01, indel_rate=0.001)
    want = _expected(data, off, 31, 19, 128)
    with gpu_ctx.job(data, off, 31, 19, 128) as job:
        job.phase1()
        r = job.autotune(sample_occurrences=200000)
        assert r["trials"] >= 20 and r["best_ms"] > 0 and r["default_ms"] > 0
        assert r["best_ms"] <= r["default_ms"]
        t = job.get_tuning()
        assert t.as_dict() == r["best"]
        job.phase2()
        _check(job, want)


def test_invalid_tuning_rejected(gpu_ctx):
    data, off = G.synth(seed=1, num_colors=2, genome_len=3000)
    with gpu_ctx.job(data, off, 31, 19, 8) as job:
        for bad in ({"digit_bits": 12}, {"cell_target": 8}, {"set_slack": 1.0}, {"persistent": 2}, {"graph": 3}):
            with pytest.raises(RuntimeError, match="tuning"):
                job.set_tuning(**bad)
        assert job.get_tuning().as_dict() == G.Tuning().as_dict()


def test_context_tuning_is_inherited(gpu_ctx):
    data, off = G.synth(seed=1, num_colors=2, genome_len=3000)
    gpu_ctx.set_tuning(persistent=1, cell_target=500)
    try:
        with gpu_ctx.job(data, off, 31, 19, 8) as job:
            t = job.get_tuning()
            assert t.persistent == 1 and t.cell_target == 500
    finally:
        gpu_ctx.set_tuning()
    with gpu_ctx.job(data, off, 31, 19, 8) as job:
        assert job.get_tuning().persistent == 0


@pytest.mark.parametrize("chunk", [0, 1])
def test_widest_digit_on_128bit_keys(gpu_ctx, chunk):
    """One bucket of ~5e5 128-bit keys sorts its top 11 bits in one global
    pass; with 11-bit digits that pass would need more LDS than a workgroup
    has, so the digit narrows to 10 bits (as with one bucket per chunk after
    the tuner chose 11 bits on a multi-bucket sample)."""
    data, off = G.synth(seed=29, num_colors=5, genome_len=100000, snp_rate=0.01, indel_rate=0.001)
    k, m, B = 63, 27, 1
    want = _expected(data, off, k, m, B)
    with gpu_ctx.job(data, off, k, m, B) as job:
        job.set_tuning(digit_bits=11)
        if chunk:
            job.set_chunk_tuples(chunk)
        job.run()
        assert job.stats()["num_occ"] > 358 * 1024  # top 11 bits of the bucket take the global pass
        _check(job, want)

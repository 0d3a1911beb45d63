"""Served mode (cdbg_gpu_server_*): jobs of different sizes and k submitted
together run on several slots at once, and each job's artifacts are the same
bytes as those of the same job run alone on the plain path
(cdbg_gpu_job_run), whether the server reads the inputs in place or copies
them; a budget that admits one job at a time gives the same results.  The HBM model is checked on the CPU: monotone in the input and
bounded by chunking."""
import numpy as np
import pytest

from conftest import cdbg_gpu as G


def test_hbm_estimate_is_monotone_and_chunk_bounded():
    e = [G.hbm_estimate(n) for n in (10**6, 10**7, 10**8)]
    assert e[0] < e[1] < e[2]
    assert G.hbm_estimate(10**8, 31, 21, 10**6) < G.hbm_estimate(10**8)
    # phase 1 is not chunked: the estimate keeps growing with the input
    assert G.hbm_estimate(10**9, 31, 21, 10**6) > G.hbm_estimate(10**8, 31, 21, 10**6)
    # 128-bit k-mers need more than 64-bit ones
    assert G.hbm_estimate(10**8, 63, 27) > G.hbm_estimate(10**8, 31, 21)
    # minimizers closer to k: shorter windows, more super-k-mers and occurrences
    assert G.hbm_estimate(10**8, 31, 27) > G.hbm_estimate(10**8, 31, 21) > G.hbm_estimate(10**8, 31, 11)
    assert G.hbm_estimate(10**8, 31, 31) == 0


JOBS = [  # (seed, colours, genome, k, m, buckets)
    (1, 3, 30_000, 31, 19, 64),
    (2, 6, 12_000, 63, 27, 16),
    (3, 2, 60_000, 31, 21, 256),
    (4, 5, 20_000, 21, 11, 32),
    (5, 4, 40_000, 63, 31, 128),
    (6, 8, 8_000, 31, 15, 8),
]


def _inputs():
    out = []
    for seed, nc, glen, k, m, B in JOBS:
        data, off = G.synth(seed=seed, num_colors=nc, genome_len=glen, snp_rate=0.01, indel_rate=0.001,
                            pav_frac=0.02, pav_min=200, pav_max=800)
        out.append((data, off, k, m, B))
    return out


def _alone(ctx, data, off, k, m, B):
    with ctx.job(data, off, k, m, B) as j:
        j.run()
        w, o = j.color_sets()
        return j.stats(), j.frags(), w, o


@pytest.mark.gpu
@pytest.mark.parametrize("slots,budget,borrow", [(3, 0, True), (2, 1, False)])
def test_served_results_equal_alone(gpu_ctx, slots, budget, borrow):
    inputs = _inputs()
    want = [_alone(gpu_ctx, *x) for x in inputs]
    if budget == 1:  # room for the largest job only: admission runs them one at a time
        budget = max(G.hbm_estimate(len(x[0]), x[2], x[3]) for x in inputs) + (1 << 20)
    with gpu_ctx.server(slots=slots, hbm_budget=budget, borrow_inputs=borrow) as srv:
        ids = [srv.submit(*x) for x in inputs]
        assert len(srv._held) == len(ids)
        for jid, (st, frags, words, offs) in zip(ids, want):
            s = srv.poll(jid, wait=True)
            assert s["state"] == 2
            assert s["submit_s"] <= s["start_s"] <= s["done_s"]
            assert s["attempts"] == 1
            for key in ("num_kmers", "num_color_sets", "num_frags", "frag_bytes", "cs_bits", "num_nodes"):
                assert s["stats"][key] == st[key], key
            f2, w2, o2 = srv.result(jid)
            assert f2.tobytes() == frags.tobytes()
            assert np.array_equal(w2, words)
            assert np.array_equal(o2, offs)
            srv.release(jid)
        assert not srv._held  # no input array outlives its job


@pytest.mark.gpu
def test_served_job_out_of_memory_runs_again_alone(gpu_ctx):
    """A job whose device allocation fails (injected from the host) runs
    again once the other slots have handed their buffers back, and its
    results are those of the job run alone; jobs after it are unaffected."""
    inputs = _inputs()[:4]
    want = [_alone(gpu_ctx, *x) for x in inputs]
    with gpu_ctx.server(slots=2) as srv:
        G.lib().cdbg_gpu_inject_oom(40)  # inside the first jobs' phase 1 or 2
        try:
            ids = [srv.submit(*x) for x in inputs]
            attempts = []
            for jid, (st, frags, words, offs) in zip(ids, want):
                s = srv.poll(jid, wait=True)
                assert s["state"] == 2
                attempts.append(s["attempts"])
                assert s["stats"]["num_kmers"] == st["num_kmers"]
                f2, w2, o2 = srv.result(jid)
                assert f2.tobytes() == frags.tobytes()
                assert np.array_equal(w2, words)
                assert np.array_equal(o2, offs)
                srv.release(jid)
            assert sorted(attempts) == [1, 1, 1, 2]  # exactly one job met the injected failure
        finally:
            G.lib().cdbg_gpu_inject_oom(-1)
    a = G.alloc_stats()
    assert a["trims"] >= 1 or a["cached_bytes"] == 0


@pytest.mark.gpu
def test_served_rejects_bad_requests(gpu_ctx):
    data, off = G.synth(seed=9, num_colors=2, genome_len=5000)
    with gpu_ctx.server(slots=2) as srv:
        with pytest.raises(RuntimeError, match="m must satisfy"):
            srv.submit(data, off, 31, 31, 16)
        with pytest.raises(RuntimeError, match="non-decreasing"):
            srv.submit(data, off[::-1].copy(), 31, 15, 16)
        with pytest.raises(RuntimeError, match="unknown job id"):
            srv.poll(12345)
        jid = srv.submit(data, off, 31, 15, 16)
        assert srv.poll(jid, wait=True)["state"] == 2
        srv.release(jid)
        with pytest.raises(RuntimeError, match="unknown job id"):
            srv.result(jid)

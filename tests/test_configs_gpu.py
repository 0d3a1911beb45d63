"""GPU parity at the benchmark workloads (the reference's semantics at -t 1,
byte for byte against the oracle on the same inputs):

* small plain build: 10 colours x 200 kbp (seed 1, 1% SNPs, 0.1% indels of
  1-10 bp), uncompressed 80-column FASTA files through cdbg-build -k 31 -m 19
  -b 6 --keep-tmp, against the oracle's phase 2 over the bucket files the run
  kept and the oracle's phases 1+2 over the same sequences; the final .fa,
  .u2c, .color_sets and metadata against the oracle's fragments stitched;
* pangenome parameters (k=31, m=21, 1024 buckets, 2% SNPs, 0.2% indels, 1%
  presence/absence deletions, 50 inversions) on the first 30 colours of the
  5 Mbp base genome;
* skewed buckets: 20 colours x 2 Mbp (seed 3), k=63 (128-bit k-mers), m=27,
  4096 buckets with Zipf(1.1) shares of the minimizer space;
* serving: the first 40 requests of the serving stream (colours log-uniform in
  [2, 200], genomes 0.5-5 Mbp, k in {31, 63}) on 2 slots: every result equals
  the same request run alone on the plain path; every tenth request (those
  with index 2 mod 10, the ones small enough for the oracle to finish in
  seconds) also equals the oracle.
"""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

import ccdbg_check as C
import oracle as O
from conftest import PKG_DIR, ROOT, cdbg_gpu as G
from test_cli import check_final_outputs

sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the serving stream and the pangenome parameters)

pytestmark = pytest.mark.gpu
CLI = os.path.join(PKG_DIR, "bin", "cdbg-build")
THREADS = 16


def _assert_job(ob, s, frags, words, offs):
    assert s["num_kmers"] == ob.stat("num_kmers")
    assert s["num_nodes"] == ob.stat("num_nodes")
    assert s["num_frags"] == ob.stat("num_frags")
    assert s["num_color_sets"] == ob.stat("num_sets")
    assert s["cs_bits"] == ob.stat("cs_bits")
    assert np.array_equal(frags, ob.frags()), "frag_unitigs.bin bytes differ"
    ow, oo = ob.color_sets()
    assert np.array_equal(offs, oo), "colour-set offsets differ"
    assert np.array_equal(words, ow), "colour-set words differ"


def _gpu(ctx, data, off, k, m, B, cuts=None):
    with ctx.job(data, off, k, m, B) as job:
        if cuts is not None:
            job.set_bucket_map(cuts)
        job.run()
        w, o = job.color_sets()
        return job.stats(), job.frags(), w, o


@pytest.mark.timeout(300)
def test_small_plain_build_through_cli(gpu_ctx, tmp_path):
    data, off = G.synth(seed=1, num_colors=10, genome_len=200_000, snp_rate=0.01, indel_rate=0.001,
                        indel_max=10, threads=THREADS)
    paths = []
    for c in range(10):
        seq = bytes(data[off[c]:off[c + 1] - 1])
        p = tmp_path / f"color_{c}.fa"
        p.write_bytes(b">color_%d\n" % c + b"\n".join(seq[i:i + 80] for i in range(0, len(seq), 80)) + b"\n")
        paths.append(str(p))
    lst = tmp_path / "list.txt"
    lst.write_text("\n".join(paths) + "\n")
    d, out = tmp_path / "tmp", tmp_path / "out"
    r = subprocess.run([CLI, "-i", str(lst), "-o", str(out), "-k", "31", "-m", "19", "-b", "6", "-t", "8",
                        "-d", str(d), "--keep-tmp"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    ob = O.Build.from_bucket_files(str(d), 64, 31, 10, threads=THREADS)
    assert (d / "frag_unitigs.bin").read_bytes() == ob.frags().tobytes()
    n = check_final_outputs(out, d, ob, 31, 10)
    # the unitig count of an independent test-side stitch over the same fragments
    assert n == len(C.split_monochromatic(C.stitch(C.parse_frags(ob.frags()), 31), 31))
    # the same sequences in memory: phases 1+2 of the oracle and of the GPU
    om = O.Build(data, off, 31, 19, 64, threads=THREADS)
    assert om.stat("num_kmers") == ob.stat("num_kmers")
    _assert_job(om, *_gpu(gpu_ctx, data, off, 31, 19, 64))


@pytest.mark.timeout(300)
def test_pangenome_parameters_30_colours(gpu_ctx):
    data, off = G.synth(num_colors=30, genome_len=5_000_000, threads=THREADS, **bench.WORKLOAD)
    ob = O.Build(data, off, 31, 21, 1024, threads=THREADS)
    _assert_job(ob, *_gpu(gpu_ctx, data, off, 31, 21, 1024))


@pytest.mark.timeout(300)
def test_skewed_buckets_k63_4096_zipf(gpu_ctx):
    c = bench.BATCHED
    data, off = G.synth(seed=3, num_colors=20, genome_len=c["genome"], snp_rate=0.01, indel_rate=0.001,
                        indel_max=10, threads=THREADS)
    cuts = G.zipf_bucket_cuts(c["buckets"], c["zipf"])
    ob = O.Build(data, off, c["k"], c["m"], c["buckets"], threads=THREADS, cuts=cuts)
    got = _gpu(gpu_ctx, data, off, c["k"], c["m"], c["buckets"], cuts=cuts)
    _assert_job(ob, *got)


def _digest(frags, words, offs):
    h = hashlib.blake2b(digest_size=16)
    for a in (frags, words, offs):
        h.update(np.ascontiguousarray(a).view(np.uint8).data)
    return h.hexdigest()


@pytest.mark.timeout(600)
def test_serving_stream_40_requests(gpu_ctx):
    jobs, _ = bench.serve_workload(40, 2.0, seed=5)
    inputs = []
    for j in jobs:
        data, off = G.synth(seed=j["seed"], num_colors=j["num_colors"], genome_len=j["genome_len"],
                            snp_rate=j["snp_rate"], indel_rate=j["indel_rate"], threads=THREADS)
        inputs.append((data, off, j["k"], j["m"], j["B"]))
    want = []
    for i, x in enumerate(inputs):
        s, f, w, o = _gpu(gpu_ctx, *x)
        want.append((s["num_kmers"], s["num_color_sets"], _digest(f, w, o)))
        if i % 10 == 2:
            _assert_job(O.Build(*x, threads=THREADS), s, f, w, o)
    with gpu_ctx.server(slots=2) as srv:
        ids = [srv.submit(*x) for x in inputs]
        for i, jid in enumerate(ids):
            st = srv.poll(jid, wait=True)
            assert st["state"] == 2, (i, st)
            f, w, o = srv.result(jid)
            got = (st["stats"]["num_kmers"], st["stats"]["num_color_sets"], _digest(f, w, o))
            assert got == want[i], f"request {i} differs between served and alone"
            srv.release(jid)

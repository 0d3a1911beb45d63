#!/usr/bin/env python3
"""Phases 1+2 (bucket-write + bucket-process) throughput on one MI355X.

Workload: the bacterial-pangenome-shaped build -- 500 colours derived from a
5 Mbp uniform-random base genome, each with 2% SNPs, 0.2% indels (1-10 bp),
1% of bases removed as 1-5 kbp presence/absence deletions and 50 inversions
of 10-100 kbp; k=31 (64-bit k-mers), m=21, 2^10 buckets.  Sequences are
generated with the seeded generator and staged in HBM before timing (the
input files' gzip decode is host work outside the timed region).

A step is one full pass of phase 1 and phase 2 over the resident input,
producing the frag_unitigs.bin stream and the colour-set stream on the
device.  value = distinct k-mers (metadata num_kmers) x steps / wall time.

Printed JSON also carries the roofline of the dominant kernel (algorithmic
bytes per launch / its average HIP-event duration on the launch stream) and
a CPU baseline: the oracle (a multi-threaded CPU restatement of the same
phases) timed on a bounded sample of the same workload.
"""
import argparse
import importlib.util
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "cdbg_gpu", os.path.join(ROOT, "cdbg-builder_amd", "cdbg_gpu.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cdbg_gpu"] = mod
    spec.loader.exec_module(mod)
    return mod


WORKLOAD = dict(seed=2, snp_rate=0.02, indel_rate=0.002, indel_max=10, pav_frac=0.01,
                pav_min=1000, pav_max=5000, n_inversions=50, inv_min=10000, inv_max=100000)


def record_bytes(s, k):
    """Uncompressed bucket-record bytes R of the job: per super-k-mer
    varint(1 colour) + varint(colour) + varint(len << 4 | flags) + ceil(len/4)
    packed bases (super_kmer.hpp:105-117); colours and lengths below 2^14
    take 2-byte varints, the length sum is occurrences + (k - 1) per record."""
    ns, n = s["num_supers"], s["num_tuples"]
    return ns * 5 + (n + ns * (k - 1)) // 4 + ns // 2


def step_bytes(name, s, k, launches, chunks):
    """Algorithmic (compulsory) HBM bytes of all launches of `name` in one
    step, by the per-kernel formulas of the hot-path scope table: tuples of
    W_k + T bytes (8 + 4 with 64-bit k-mers, 16 + 4 with 128-bit ones), every
    buffer touched once.  Kernels without a formula return None (their time
    still counts in the phase-2 aggregate).  `launches` is the kernel's launch
    count per step, `chunks` the phase-2 chunks."""
    wk = 8 if k <= 31 else 16
    t = wk + 4
    n, nn, ns, G = s["num_tuples"], s["num_nodes"], s["num_supers"], s["num_bases"]
    F, nd, sets, ti = s["frag_bytes"], s["num_kmers"], s["num_color_sets"], s["total_integers"]
    R = record_bytes(s, k)
    passes = launches / max(chunks, 1)  # global radix passes per chunk
    table = {
        "k_pack2bit": G + G // 4,                         # A + G/4
        "k_superkmers": G // 4 + 16 * ns,                 # G/4 + 16 N_sk
        "k_radix_scatter:buckets": 16 * ns + 2 * R,       # bucket scatter: 16 N_sk + 2R
        "k_expand": R + n * t,                            # R + N_occ (W_k + T)
        "k_radix_scatter:tuples": passes * 2 * n * t,     # each radix pass: 2 N_occ (W_k + T)
        "k_cell_sort": 2 * n * t,                         # the last pass, in LDS: 2 N_occ (W_k + T)
        "k_nodes": n * t + nn * (wk + 5),                 # segment reduce: N_occ (W_k + T) + N_b (W_k + 5)
        "k_set_insert": 4 * ti + 4 * nd + 32 * sets,      # colour-set hash/intern
        "k_dir_build": nn * (wk + 4),                     # table build: N_b (W_k + 4)
        "k_masks": 8 * (wk + 4) * nn + nn,                # degree: 8 (W_k + 4) N_b + N_b
        "k_frag_write": nn * (wk + 4) + F,                # fragment emit: N_b (W_k + 4) + F
    }
    return table.get(name)


PHASE1_KERNELS = ("k_pack2bit", "k_superkmers", "k_zip_supers", "k_split_supers", "k_bucket_offsets",
                  "k_radix_hist:buckets", "k_radix_scatter:buckets", "k_count_kmers")


def roofline(kt, s, k, steps, chunks, ms_step):
    """Roofline of the dominant kernel that has an algorithmic-byte formula
    (bytes per launch / average launch time, both from `steps` timed steps),
    the phase-2 aggregate (modelled bytes of every phase-2 kernel over the
    time of all phase-2 kernels, modelled or not), and the six costliest
    kernels in ms per step."""
    roof = None
    agg_b = agg_ms = 0.0
    for name, (cnt, ms) in sorted(kt.items(), key=lambda x: -x[1][1]):
        b = step_bytes(name, s, k, cnt / steps, chunks) if cnt else None
        if name not in PHASE1_KERNELS:
            agg_ms += ms
            agg_b += (b or 0) * steps
        if b is None or roof is not None:
            continue
        per_launch = b * steps / cnt
        avg_ms = ms / cnt
        ach = per_launch / (avg_ms / 1e3) / 1e9
        roof = {"bound": "hbm", "kernel": name, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": int(per_launch), "avg_launch_ms": round(avg_ms, 3),
                "launches_per_step": cnt // steps,
                "share_of_step": round(ms / steps / ms_step, 3)}
    if roof is not None and agg_ms > 0:
        ach = agg_b / (agg_ms / 1e3) / 1e9
        roof["phase2_modelled"] = {"achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
                                   "share_of_step": round(agg_ms / steps / ms_step, 3)}
    top = sorted(kt.items(), key=lambda x: -x[1][1])[:6]
    return roof, {n: round(ms / steps, 2) for n, (c, ms) in top}


def kernel_base(name):
    n = name.replace("(anonymous namespace)::", "").split("(")[0]
    return n.split("<")[0].split("::")[-1]


def pmc_child(args):
    """One phases-1+2 run of the workload and a 1 GiB copy probe, for a
    profiler run of the parent (no output besides the counters)."""
    G = load_pkg()
    data, off = G.synth(num_colors=args.colors, genome_len=args.genome, threads=args.cpu_cores, **WORKLOAD)
    ctx = G.Context(0)
    with ctx.job(data, off, args.k, args.m, args.buckets) as job:
        job.run()
    ctx.copy_probe(1 << 30, 1)
    ctx.close()


def measure_traffic(args, kernel, launches):
    """HBM bytes per launch of `kernel` from two rocprofv3 counter runs of
    one step (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by
    what the counter reads for the 1 GiB copy probe of the same run
    (k_copy_probe reads and writes 1 GiB).  None when the profiler is absent
    or a run fails."""
    import csv
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof or not kernel:
        return None
    base = kernel.split(":")[0]
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        with tempfile.TemporaryDirectory(dir="/tmp") as d:
            cmd = [prof, "--pmc", counter, "-d", d, "-o", "run", "--output-format", "csv", "--",
                   sys.executable, os.path.abspath(__file__), "--pmc-child", "--colors", str(args.colors),
                   "--genome", str(args.genome), "--k", str(args.k), "--m", str(args.m),
                   "--buckets", str(args.buckets), "--cpu-cores", str(args.cpu_cores)]
            try:
                subprocess.run(cmd, check=True, timeout=600, stdout=subprocess.DEVNULL,
                               stderr=subprocess.DEVNULL, env=dict(os.environ, TMPDIR="/tmp"))
            except (subprocess.SubprocessError, OSError):
                return None
            rows = []
            for root, _, files in os.walk(d):
                for f in files:
                    if f.endswith("counter_collection.csv"):
                        rows += list(csv.DictReader(open(os.path.join(root, f))))
            kv = [float(r["Counter_Value"]) for r in rows if kernel_base(r["Kernel_Name"]) == base]
            probe = [float(r["Counter_Value"]) for r in rows if kernel_base(r["Kernel_Name"]) == "k_copy_probe"]
            if not kv or not probe:
                return None
            scale = float(1 << 30) / (probe[-1] * 1024.0)  # true bytes per counted byte
            # the step's bytes of the kernel (a tagged kernel shares its name with
            # other sorts: the largest `launches` dispatches are its own) over the
            # timed run's launches per step: the child may plan other chunks
            # (its free-memory reading differs under the profiler), and the
            # bytes of a step do not depend on the chunking
            kv = sorted(kv)[-launches:] if base != kernel else kv
            out[counter] = sum(kv) * 1024.0 * scale / launches
    return {"fetch_bytes": int(out["FETCH_SIZE"]), "write_bytes": int(out["WRITE_SIZE"]),
            "traffic": int(out["FETCH_SIZE"] + out["WRITE_SIZE"])}


# Skewed-bucket (batched) workload: 200 colours of 2 Mbp genomes (1% SNPs,
# 0.1% indels of 1-10 bp), k=63 (128-bit k-mers), m=27, 4096 buckets whose
# shares of the minimizer space follow Zipf(1.1).
BATCHED = dict(colors=200, genome=2_000_000, k=63, m=27, buckets=4096, zipf=1.1)


def batched_leg(G, ctx, steps, warmup, threads):
    """Phases 1+2 over thousands of buckets of skewed sizes (one launch per
    kernel covers every bucket of a chunk): distinct k-mers/s, bucket skew,
    roofline of its dominant kernel."""
    import numpy as np
    c = BATCHED
    t_gen = time.perf_counter()
    data, off = G.synth(seed=3, num_colors=c["colors"], genome_len=c["genome"], snp_rate=0.01,
                        indel_rate=0.001, indel_max=10, threads=threads)
    print(f"bench: batched workload {c['colors']} colours, {len(data) / 1e9:.2f} GB in "
          f"{time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)
    with ctx.job(data, off, c["k"], c["m"], c["buckets"]) as job:
        job.set_bucket_map(G.zipf_bucket_cuts(c["buckets"], c["zipf"]))
        for _ in range(warmup):
            job.run()
        # the same workload with the default knobs (one workgroup per cell),
        # then tuned by the autotuner over a sample of its first buckets
        t0 = time.perf_counter()
        for _ in range(steps):
            job.run()
        plain_elapsed = time.perf_counter() - t0
        job.phase1()
        t0 = time.perf_counter()
        tune = job.autotune()
        tune["tuning_s"] = round(time.perf_counter() - t0, 3)
        job.run()
        job.set_profiling(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            job.run()
        elapsed = time.perf_counter() - t0
        s = job.stats()
        kt = job.kernel_times()
        chunks = job.num_chunks()
        sizes = np.diff(job.bucket_offsets().astype(np.int64))
        # the same buckets one at a time: a chunk per bucket, so every kernel
        # is launched (and every count read back) per bucket, as a GPU port
        # of the reference's per-bucket scheduler (bucket_walker.hpp:643-779)
        # would run them
        job.set_profiling(False)
        job.set_chunk_tuples(1)
        t0 = time.perf_counter()
        job.run()
        pb_elapsed = time.perf_counter() - t0
        pb_chunks = job.num_chunks()
        if job.stats()["num_kmers"] != s["num_kmers"]:
            raise RuntimeError("per-bucket run differs from the batched run")
        job.set_chunk_tuples(0)
    pb_value = s["num_kmers"] / pb_elapsed
    plain_value = s["num_kmers"] * steps / plain_elapsed
    ms_step = 1e3 * elapsed / steps
    roof, kernels = roofline(kt, s, c["k"], steps, chunks, ms_step)
    value = s["num_kmers"] * steps / elapsed
    print(f"bench: batched {steps} steps in {elapsed:.2f} s, {value / 1e9:.3f} G k-mers/s", file=sys.stderr,
          flush=True)
    return {"metric": "distinct k-mers/s (phases 1+2)", "value": round(value, 1), "ms_per_step": round(ms_step, 3),
            "default_knobs_value": round(plain_value, 1), "tuned_over_default": round(value / plain_value, 3),
            "per_bucket": {"value": round(pb_value, 1), "ms_per_step": round(1e3 * pb_elapsed, 1),
                           "chunks": pb_chunks},
            "batched_over_per_bucket": round(value / pb_value, 3),
            "autotune": tune,
            "steps": steps, "ms_phase1": round(s["ms_phase1"], 2), "ms_phase2": round(s["ms_phase2"], 2),
            "chunks": chunks, "num_kmers": s["num_kmers"], "num_occ": s["num_occ"],
            "bucket_supers": {"max": int(sizes.max()), "median": int(np.median(sizes)), "min": int(sizes.min())},
            "roofline": roof, "top_kernels_ms_per_step": kernels,
            "workload": f"{c['colors']} colours x {c['genome'] / 1e6:g} Mbp, 1% SNP, 0.1% indel; "
                        f"k={c['k']} (128-bit k-mers) m={c['m']}, {c['buckets']} buckets with "
                        f"Zipf({c['zipf']}) shares of the minimizer space"}


# Large resident build: colours of 5 Mbp genomes at 1.5% divergence (1.5%
# SNPs, 0.15% indels of 1-10 bp), k=31, m=21, 2^14 buckets; phase 1 runs in
# colour ranges and phase 2 in bucket chunks, both sized from free HBM.
LARGE = dict(genome=5_000_000, k=31, m=21, buckets=1 << 14, snp_rate=0.015, indel_rate=0.0015)


def large_leg(G, ctx, colors, steps, threads):
    """Phases 1+2 of the large resident build at `colors` colours: distinct
    k-mers/s, phase-1 colour ranges, phase-2 chunks, device memory."""
    c = LARGE
    t_gen = time.perf_counter()
    data, off = G.synth(seed=4, num_colors=colors, genome_len=c["genome"], snp_rate=c["snp_rate"],
                        indel_rate=c["indel_rate"], indel_max=10, threads=threads)
    print(f"bench: large workload {colors} colours, {len(data) / 1e9:.2f} GB in "
          f"{time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)
    with ctx.job(data, off, c["k"], c["m"], c["buckets"]) as job:
        del data
        job.phase1()
        print(f"bench: large phase 1 done: {job.num_phase1_ranges()} colour ranges, "
              f"{G.lib().cdbg_gpu_device_bytes() / 2**30:.1f} GiB on device", file=sys.stderr, flush=True)
        job.phase2()
        peak = G.lib().cdbg_gpu_device_bytes()
        print(f"bench: large warmup done ({peak / 2**30:.1f} GiB on device)", file=sys.stderr, flush=True)
        job.set_profiling(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            job.run()
        elapsed = time.perf_counter() - t0
        s = job.stats()
        kt = job.kernel_times()
        top = sorted(kt.items(), key=lambda x: -x[1][1])[:12]
        res = {"metric": "distinct k-mers/s (phases 1+2)", "value": round(s["num_kmers"] * steps / elapsed, 1),
               "top_kernels_ms_per_step": {n: round(ms / steps, 1) for n, (c, ms) in top},
               "ms_per_step": round(1e3 * elapsed / steps, 1), "steps": steps,
               "ms_phase1": round(s["ms_phase1"], 1), "ms_phase2": round(s["ms_phase2"], 1),
               "phase1_ranges": job.num_phase1_ranges(), "phase2_chunks": job.num_chunks(),
               "num_bases": s["num_bases"], "num_kmers": s["num_kmers"], "num_occ": s["num_occ"],
               "device_gib_after_run": round(peak / 2**30, 1),
               "workload": f"{colors} colours x {c['genome'] / 1e6:g} Mbp, 1.5% SNP, 0.15% indel; k={c['k']} "
                           f"m={c['m']} B={c['buckets']}; input resident in HBM"}
    print(f"bench: large {json.dumps(res)}", file=sys.stderr, flush=True)
    return res


def serve_workload(n, rate, seed):
    """Build requests of the serving configuration: colours log-uniform in
    [2, 200], genomes U[0.5, 5] Mbp at U[0.5, 3]% divergence (SNPs; indels at a
    tenth of it), k drawn from {31, 63}; Poisson arrivals at `rate` jobs/s."""
    import numpy as np
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        nc = int(round(math.exp(rng.uniform(math.log(2), math.log(200)))))
        glen = int(rng.uniform(0.5e6, 5e6))
        div = float(rng.uniform(0.005, 0.03))
        k = 31 if rng.random() < 0.5 else 63
        jobs.append(dict(seed=10_000 + i, num_colors=nc, genome_len=glen, snp_rate=div, indel_rate=div / 10,
                         k=k, m=21 if k == 31 else 27, B=512))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    return jobs, arrivals


def served_leg(G, ctx, n_jobs, slots, rate, threads):
    """Plain (a new job per request, one at a time) against served (resident
    per-slot buffers; one slot, then `slots` jobs overlapping on the device)
    over the same requests: aggregate distinct k-mers/s when every request is
    queued at once, and p50/p99 job latency (submit -> artifacts complete in
    device memory) under Poisson arrivals; the plain latencies replay the arrivals through a
    FIFO queue with the measured plain service times."""
    import numpy as np
    jobs, arrivals = serve_workload(n_jobs, rate, seed=5)
    t_gen = time.perf_counter()
    inputs = []
    for j in jobs:
        data, off = G.synth(seed=j["seed"], num_colors=j["num_colors"], genome_len=j["genome_len"],
                            snp_rate=j["snp_rate"], indel_rate=j["indel_rate"], threads=threads)
        inputs.append((data, off, j["k"], j["m"], j["B"]))
    bases = sum(len(x[0]) for x in inputs)
    print(f"bench: serving workload {n_jobs} jobs, {bases / 1e9:.2f} Gbp, generated in "
          f"{time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)

    def plain(x):  # the artifacts stay in device memory, as the served jobs' do
        with ctx.job(*x) as jb:
            jb.run()
            return jb.stats()["num_kmers"]

    for x in inputs[:2]:  # both key widths: load the kernels once
        plain(x)
    service, kmers = [], []
    for x in inputs:
        t0 = time.perf_counter()
        kmers.append(plain(x))
        service.append(time.perf_counter() - t0)
    plain_tput = sum(kmers) / sum(service)
    # how busy the plain path keeps the GPU: kernel time (HIP events around
    # every launch, so an upper bound) over the plain wall time; served
    # overlap can only fill the rest, so about 1 / busy bounds served / plain
    kernel_s = 0.0
    for x in inputs:
        with ctx.job(*x) as jb:
            jb.set_profiling(True)
            jb.run()
            kernel_s += sum(ms for _, ms in jb.kernel_times().values()) / 1e3
    plain_busy = kernel_s / sum(service)
    fin, lat_plain = 0.0, []
    for a, sv in zip(arrivals, service):
        fin = max(fin, a) + sv
        lat_plain.append(fin - a)
    print(f"bench: plain {sum(service):.2f} s for {n_jobs} jobs", file=sys.stderr, flush=True)

    def alloc_delta(a, reruns):
        b = G.alloc_stats()
        return {"reruns_alone": reruns, "trims": b["trims"] - a["trims"], "mallocs": b["mallocs"] - a["mallocs"],
                "malloc_gb": round((b["malloc_bytes"] - a["malloc_bytes"]) / 1e9, 2),
                "malloc_ms": round((b["malloc_ns"] - a["malloc_ns"]) / 1e6, 1),
                "trim_ms": round((b["trim_ns"] - a["trim_ns"]) / 1e6, 1)}

    def closed_batch(srv, n_slots):
        warm = [srv.submit(*x) for x in inputs[:n_slots]]  # every slot loads its buffers once
        for jid in warm:
            srv.poll(jid, wait=True)
            srv.release(jid)
        a = G.alloc_stats()
        t0 = time.perf_counter()
        ids = [srv.submit(*x) for x in inputs]
        got, reruns = [], 0
        for jid in ids:
            st = srv.poll(jid, wait=True)
            got.append(st["stats"]["num_kmers"])
            reruns += st["attempts"] > 1
            srv.release(jid)
        wall = time.perf_counter() - t0
        if got != kmers:
            raise RuntimeError("served results differ from the plain path")
        alloc = alloc_delta(a, reruns)
        print(f"bench: served, {n_slots} slot(s), all queued: {wall:.2f} s, allocator {alloc}",
              file=sys.stderr, flush=True)
        return sum(got) / wall, alloc

    with ctx.server(slots=1) as srv:
        resident_tput, _ = closed_batch(srv, 1)
    with ctx.server(slots=slots) as srv:
        served_tput, alloc_closed = closed_batch(srv, slots)
        a0 = G.alloc_stats()
        t0 = time.perf_counter()
        ids = []
        for a, x in zip(arrivals, inputs):
            dt = t0 + a - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(*x))
        lat_served, reruns = [], 0
        for jid in ids:
            st = srv.poll(jid, wait=True)
            lat_served.append(st["done_s"] - st["submit_s"])
            reruns += st["attempts"] > 1
            srv.release(jid)
        alloc_open = alloc_delta(a0, reruns)
        print(f"bench: served, Poisson arrivals: allocator {alloc_open}", file=sys.stderr, flush=True)
    pct = lambda v, q: round(1e3 * float(np.percentile(v, q)), 1)
    return {"jobs": n_jobs, "slots": slots, "bases": bases, "distinct_kmers": sum(kmers),
            "plain": {"value": round(plain_tput, 1), "p50_ms": pct(lat_plain, 50), "p99_ms": pct(lat_plain, 99),
                      "gpu_busy": round(plain_busy, 3)},
            "served_1_slot": {"value": round(resident_tput, 1)},
            "served": {"value": round(served_tput, 1), "p50_ms": pct(lat_served, 50),
                       "p99_ms": pct(lat_served, 99), "allocator_all_queued": alloc_closed,
                       "allocator_poisson": alloc_open},
            "served_over_plain": round(served_tput / plain_tput, 3),
            "arrival_rate": rate,
            "workload": f"{n_jobs} build requests: colours log-uniform [2,200], genomes U[0.5,5] Mbp, "
                        f"divergence U[0.5,3]%, k in {{31,63}} (m 21/27), B=512; throughput with every "
                        f"request queued at once, latency under Poisson arrivals at {rate} jobs/s"}


def cpu_baseline(G, sample_colors, genome, k, m, B, cores):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    data, off = G.synth(num_colors=sample_colors, genome_len=genome, threads=cores, **WORKLOAD)
    ob = O.Build(data, off, k, m, B, threads=cores)
    t1, t2 = ob.times()
    nk, nocc = ob.stat("num_kmers"), ob.stat("num_occ")
    per_thread = nocc / (t1 + t2) / cores
    return {"value": nk / (t1 + t2), "unit": "distinct k-mers/s", "cores": cores, "kind": "port",
            "occurrences_per_s": round(nocc / (t1 + t2), 1), "bases_per_s": round(len(data) / (t1 + t2), 1),
            "phase_s": [round(t1, 2), round(t2, 2)],
            # the reference's own published run (README.md:165,183,186): 1.573e11
            # k-mer occurrences through phases 1+2 in 14,519.9 + 15,609.1 s on
            # 48 threads, scratch on disk
            "occurrences_per_s_per_thread": round(per_thread, 1),
            "reference_published_occurrences_per_s_per_thread": round(1.57311093967e11 / 30129.0 / 48, 1),
            "sample": f"first {sample_colors} of the workload's colours ({len(data)/1e6:.1f} Mbp), "
                      f"phases 1+2 of the oracle, {nk} distinct k-mers ({nocc} k-mer occurrences) "
                      f"in {t1 + t2:.1f} s"}


def file_path_leg(G, ctx, colors, genome, k, m, B, threads):
    """The drop-in file path, timed by wall clock: gzipped FASTA files (one
    per colour, level 6) through cdbg_gpu_bucket_write (inflate + parse,
    upload, phase 1, LZ4 bucket files) and cdbg_gpu_bucket_process (read +
    decompress bucket files, upload, phase 2, frag_unitigs.bin and
    .color_sets), then the device stitch (cdbg_gpu_stitch_file) and the host
    emit (.fa, .u2c, finished .color_sets, metadata).  The host stitch of the
    same file is timed beside it (stitch_host_s) and must give the same
    statistics and, emitted, byte-identical .fa and .u2c."""
    import shutil
    import tempfile
    d = tempfile.mkdtemp(prefix="cdbg_bench_", dir="/tmp")
    try:
        t0 = time.perf_counter()
        files = G.synth_fasta_files(os.path.join(d, "in"), colors, genome, gz=True, threads=threads,
                                    **WORKLOAD)
        t_gen = time.perf_counter() - t0
        gz_bytes = sum(os.path.getsize(f) for f in files)
        tmp, out = os.path.join(d, "tmp"), os.path.join(d, "out")
        os.makedirs(tmp)
        t = {}
        t0 = time.perf_counter()
        unc = ctx.bucket_write(files, k, m, B, tmp, threads=threads)
        t["bucket_write_s"] = time.perf_counter() - t0
        frag = os.path.join(tmp, "frag_unitigs.bin")
        t0 = time.perf_counter()
        st = ctx.bucket_process(tmp, B, k, colors, frag, out + ".color_sets", threads=threads)
        t["bucket_process_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        with ctx.stitch_file(frag, k, threads=threads) as U:
            sst = U.stats()
            nu = sst["num_unitigs"]
            t["stitch_s"] = time.perf_counter() - t0
            t0 = time.perf_counter()
            U.emit_fasta(out)
        G.finalize_color_sets(out + ".color_sets", st)
        G.write_metadata(out, k, st["num_kmers"], colors, nu, st["num_color_sets"])
        t["emit_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        with G.stitch_file(frag, k, threads=threads) as H:
            t_host = time.perf_counter() - t0
            if H.stats() != sst:
                raise RuntimeError("device and host stitch disagree")
            H.emit_fasta(out + ".host")
        import filecmp
        same_files = all(filecmp.cmp(out + ext, out + ".host" + ext, shallow=False) for ext in (".fa", ".u2c"))
        if not same_files:
            raise RuntimeError("device and host stitch wrote different .fa / .u2c")
        # bucket-process again, part by part through the job calls it is made
        # of (untimed above): the fragment stream's write into the page cache
        # is the part that varies from machine to machine
        os.sync()
        parts = {}
        t0 = time.perf_counter()
        job = ctx.job_from_bucket_files(tmp, B, k, colors, threads=threads)
        try:
            parts["read_buckets"] = time.perf_counter() - t0
            t0 = time.perf_counter()
            job.phase2()
            parts["phase2"] = time.perf_counter() - t0
            t0 = time.perf_counter()
            job.write_phase2_files(None, out + ".cs2")
            parts["color_sets"] = time.perf_counter() - t0
            t0 = time.perf_counter()
            job.write_phase2_files(frag + ".2", out + ".cs2")
            parts["frag_file_and_color_sets"] = time.perf_counter() - t0
        finally:
            job.close()
        p12 = t["bucket_write_s"] + t["bucket_process_s"]
        res = {"metric": "distinct k-mers/s (phases 1+2, files)", "value": round(st["num_kmers"] / p12, 1),
               "end_to_end_kmers_per_s": round(st["num_kmers"] / sum(t.values()), 1),
               "num_kmers": st["num_kmers"], "num_unitigs": nu, "gz_input_bytes": gz_bytes,
               "bucket_bytes_uncompressed": int(unc.sum()), "frag_bytes": os.path.getsize(frag),
               "fa_bytes": os.path.getsize(out + ".fa"), "host_threads": threads,
               **{kk: round(v, 3) for kk, v in t.items()},
               "stitch_host_s": round(t_host, 3), "fa_u2c_identical_to_host_stitch": same_files,
               "bucket_process_parts_s": {kk: round(v, 3) for kk, v in parts.items()},
               "workload": f"first {colors} colours of the pangenome workload as gzipped FASTA (generated "
                           f"in {t_gen:.1f} s, untimed); k={k} m={m} B={B}; scratch under /tmp"}
        print(f"bench: file path {json.dumps(res)}", file=sys.stderr, flush=True)
        return res
    finally:
        shutil.rmtree(d, ignore_errors=True)


def full_build_leg(G, job, s, ms_step, threads):
    """The rest of a complete build of the plain workload after its timed
    phases 1+2: the device stitch of the job's fragment stream where it lies
    in HBM (cdbg_gpu_job_stitch: no frag_unitigs.bin), then the host emit of
    .fa and .u2c into /tmp.  Wall clock; end to end = one step of phases 1+2
    plus these two."""
    import shutil
    import tempfile
    d = tempfile.mkdtemp(prefix="cdbg_full_", dir="/tmp")
    try:
        t0 = time.perf_counter()
        with job.stitch(threads=threads) as U:
            st = U.stats()
            t_stitch = time.perf_counter() - t0
            t0 = time.perf_counter()
            U.emit_fasta(os.path.join(d, "out"))
            t_emit = time.perf_counter() - t0
        if st["unitig_kmers"] != s["num_kmers"]:
            raise RuntimeError("the unitigs do not cover the distinct k-mers")
        total = ms_step / 1e3 + t_stitch + t_emit
        res = {"stitch_s": round(t_stitch, 3), "emit_s": round(t_emit, 3), "phases12_s": round(ms_step / 1e3, 3),
               "end_to_end_s": round(total, 3), "end_to_end_kmers_per_s": round(s["num_kmers"] / total, 1),
               "num_unitigs": st["num_unitigs"], "joins": st["joins"], "cycles": st["cycles"],
               "num_fragments": st["num_fragments"], "fa_bytes": os.path.getsize(os.path.join(d, "out.fa")),
               "host_threads": threads}
        print(f"bench: full build {json.dumps(res)}", file=sys.stderr, flush=True)
        return res
    finally:
        shutil.rmtree(d, ignore_errors=True)


def bucket_io_leg(G, ctx, job, s, k, B, colors, threads):
    """Bucket files of the plain workload through the device writer
    (identical records merged, LZ4 frames compressed on the device, files
    written by host threads) and the device reader (files read by host
    threads, frames decoded and records walked on the device), then phase 2
    over what was read, whose totals must equal the resident run's.  Wall
    clock; the writer's kernels from the job profiler."""
    import shutil
    import tempfile
    d = tempfile.mkdtemp(prefix="cdbg_buckets_", dir="/tmp")
    try:
        job.phase1()
        kt0 = job.kernel_times()
        t0 = time.perf_counter()
        unc = job.write_bucket_files(d, threads)
        t_write = time.perf_counter() - t0
        kt1 = job.kernel_times()
        wk = {}
        for name, (cnt, ms) in kt1.items():
            dm = ms - kt0.get(name, (0, 0.0))[1]
            if dm > 0.05:
                wk[name] = round(dm, 2)
        file_bytes = sum(os.path.getsize(os.path.join(d, f)) for f in os.listdir(d))
        t0 = time.perf_counter()
        fj = ctx.job_from_bucket_files(d, B, k, colors, threads)
        t_read = time.perf_counter() - t0
        t0 = time.perf_counter()
        fj.phase2()
        t_p2 = time.perf_counter() - t0
        fs = fj.stats()
        fj.close()
        same = all(fs[x] == s[x] for x in ("num_kmers", "num_frags", "num_color_sets", "frag_bytes", "cs_bits"))
        res = {"records_in": s["num_supers"], "records_out": fs["num_supers"],
               "bytes_uncompressed": int(unc.sum()), "bytes_files": file_bytes,
               "write_s": round(t_write, 3), "read_s": round(t_read, 3), "phase2_from_files_s": round(t_p2, 3),
               "write_kernels_ms": dict(sorted(wk.items(), key=lambda x: -x[1])),
               "phase2_totals_equal_resident": same, "host_threads": threads,
               "workload": f"the plain workload's {B} buckets; files under /tmp"}
        print(f"bench: bucket files {json.dumps(res)}", file=sys.stderr, flush=True)
        return res
    finally:
        shutil.rmtree(d, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--colors", type=int, default=500)
    ap.add_argument("--genome", type=int, default=5_000_000)
    ap.add_argument("--k", type=int, default=31)
    ap.add_argument("--m", type=int, default=21)
    ap.add_argument("--buckets", type=int, default=1024)
    ap.add_argument("--cpu-sample-colors", type=int, default=30)
    ap.add_argument("--cpu-cores", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=40, help="requests of the served leg (0: skip it)")
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--serve-rate", type=float, default=2.0, help="Poisson arrivals per second")
    ap.add_argument("--kernel-times", action="store_true", help="print per-kernel totals to stderr")
    ap.add_argument("--no-batched", action="store_true", help="skip the skewed-bucket (batched) leg")
    ap.add_argument("--batched-only", action="store_true", help="run only the skewed-bucket (batched) leg")
    ap.add_argument("--served-only", action="store_true", help="run only the serving leg")
    ap.add_argument("--copy-probe-only", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel over BYTES (counter calibration)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two profiler runs that measure traffic")
    ap.add_argument("--file-colors", type=int, default=100,
                    help="colours of the gzipped-FASTA file-path leg (0: skip it)")
    ap.add_argument("--large-colors", type=int, default=4000,
                    help="colours of the large resident build leg (0: skip it)")
    ap.add_argument("--large-only", action="store_true", help="run only the large resident build leg")
    ap.add_argument("--file-only", action="store_true", help="run only the file-path leg")
    ap.add_argument("--no-full-build", action="store_true",
                    help="skip the stitch and emit of the plain workload's fragments")
    ap.add_argument("--no-bucket-files", action="store_true", help="skip the bucket-file write/read leg")
    ap.add_argument("--bucket-files-only", action="store_true",
                    help="time the plain workload, then only the bucket-file leg (no traffic runs)")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: runs on one GPU only (buckets would shard across GPUs by size, with one "
              "all-to-all for colour-set interning; not built), so no figure for N > 1 is produced.",
              file=sys.stderr)
        sys.exit(2)

    G = load_pkg()
    if args.large_only:
        ctx = G.Context(0)
        print(json.dumps({"large": large_leg(G, ctx, args.large_colors or 4000, args.steps, args.cpu_cores)}))
        ctx.close()
        return
    if args.file_only:
        ctx = G.Context(0)
        print(json.dumps({"files": file_path_leg(G, ctx, args.file_colors or 100, args.genome, args.k, args.m,
                                                 args.buckets, args.cpu_cores)}))
        ctx.close()
        return
    if args.served_only:
        ctx = G.Context(0)
        print(json.dumps({"served": served_leg(G, ctx, args.serve_jobs or 40, args.serve_slots, args.serve_rate,
                                               args.cpu_cores)}))
        ctx.close()
        return
    if args.batched_only:
        ctx = G.Context(0)
        print(json.dumps({"batched": batched_leg(G, ctx, args.steps, args.warmup, args.cpu_cores)}))
        ctx.close()
        return
    if args.copy_probe_only:
        ctx = G.Context(0)
        print(json.dumps({"copy_probe_bytes": args.copy_probe_only,
                          "gbps": ctx.copy_probe(args.copy_probe_only, 3)}))
        return
    t_gen = time.perf_counter()
    data, off = G.synth(num_colors=args.colors, genome_len=args.genome, threads=args.cpu_cores,
                        **WORKLOAD)
    print(f"bench: generated {args.colors} colours, {len(data)/1e9:.2f} GB in "
          f"{time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)
    ctx = G.Context(0)
    job = ctx.job(data, off, args.k, args.m, args.buckets)
    del data
    copy_gbps = ctx.copy_probe(4 << 30, 3)
    for i in range(args.warmup):
        job.run()
        print(f"bench: warmup {i} done ({G.lib().cdbg_gpu_device_bytes() / 2**30:.1f} GiB on device)",
              file=sys.stderr, flush=True)
    job.set_profiling(True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        job.run()
    elapsed = time.perf_counter() - t0
    s = job.stats()
    kt = job.kernel_times()
    ms_step = 1e3 * elapsed / args.steps
    value = s["num_kmers"] * args.steps / elapsed
    print(f"bench: {args.steps} steps in {elapsed:.2f} s, phase1 {s['ms_phase1']:.1f} ms, "
          f"phase2 {s['ms_phase2']:.1f} ms (last step)", file=sys.stderr, flush=True)
    chunks = job.num_chunks()
    if args.kernel_times:
        print("bench: sizes " + json.dumps(s), file=sys.stderr)
        for name, (cnt, ms) in sorted(kt.items(), key=lambda x: -x[1][1]):
            b = step_bytes(name, s, args.k, cnt / args.steps, chunks)
            model = (f"  {b * args.steps / cnt / 1e9:9.3f} GB/launch  {b * args.steps / (ms / 1e3) / 1e9:8.1f} GB/s"
                     if b and cnt else "")
            print(f"  {name:32s} {cnt:6d} {ms:10.2f} ms{model}", file=sys.stderr)

    roof, kernels = roofline(kt, s, args.k, args.steps, chunks, ms_step)
    roof["copy_probe_gbps"] = round(copy_gbps, 1)

    def leg(fn, *a):
        """A secondary leg, started on an empty block cache (what earlier legs
        cached would otherwise decide its chunking and allocations) and with
        the page cache's dirty data written back (the gigabytes earlier legs
        wrote would otherwise throttle this leg's writes); its failure is
        reported in its place, not raised."""
        G.lib().cdbg_gpu_trim_cache()
        os.sync()
        try:
            return fn(*a)
        except Exception as e:  # noqa: BLE001
            print(f"bench: {fn.__name__} failed: {e}", file=sys.stderr, flush=True)
            return {"error": str(e)}

    full = None
    if not args.no_full_build:
        full = leg(full_build_leg, G, job, s, ms_step, args.cpu_cores)
    bucket_files = None
    if not args.no_bucket_files:
        bucket_files = leg(bucket_io_leg, G, ctx, job, s, args.k, args.buckets, args.colors, args.cpu_cores)
    job.close()
    G.lib().cdbg_gpu_trim_cache()  # the profiler's child processes need the HBM this process cached
    if args.bucket_files_only:
        print(json.dumps({"bucket_files": bucket_files}))
        ctx.close()
        return
    if not args.no_traffic:
        try:
            tr = measure_traffic(args, roof["kernel"], roof["launches_per_step"])
        except Exception as e:  # noqa: BLE001
            print(f"bench: traffic measurement failed: {e}", file=sys.stderr, flush=True)
            tr = None
        if tr:
            roof["traffic"] = tr["traffic"]
            roof["traffic_fetch"] = tr["fetch_bytes"]
            roof["traffic_write"] = tr["write_bytes"]
            roof["traffic_over_model"] = round(tr["traffic"] / roof["bytes_per_launch"], 3)
        print(f"bench: traffic of {roof['kernel']}: {tr}", file=sys.stderr, flush=True)
    files = None
    if args.file_colors:
        files = leg(file_path_leg, G, ctx, min(args.file_colors, args.colors), args.genome, args.k, args.m,
                    args.buckets, args.cpu_cores)
    batched = None
    if not args.no_batched:
        batched = leg(batched_leg, G, ctx, args.steps, args.warmup, args.cpu_cores)
    served = None
    if args.serve_jobs:
        served = leg(served_leg, G, ctx, args.serve_jobs, args.serve_slots, args.serve_rate, args.cpu_cores)
    large = None
    if args.large_colors:
        large = leg(large_leg, G, ctx, args.large_colors, 1, args.cpu_cores)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = leg(cpu_baseline, G, args.cpu_sample_colors, args.genome, args.k, args.m, args.buckets,
                  args.cpu_cores)
    out = {
        "metric": "distinct k-mers/s (phases 1+2)",
        "value": round(value, 1),
        "unit": "distinct k-mers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64" if args.k <= 31 else "u128",
        "data": "synthetic",
        "config": {"workload": f"pangenome {args.colors} colours x {args.genome/1e6:g} Mbp, 2% SNP, "
                               f"0.2% indel, 1% PAV, 50 inversions; k={args.k} m={args.m} "
                               f"B={args.buckets}; plain single-job path, input resident in HBM",
                   "num_kmers": s["num_kmers"], "num_occ": s["num_occ"],
                   "num_color_sets": s["num_color_sets"], "num_frags": s["num_frags"]},
        "roofline": roof,
        "top_kernels_ms_per_step": kernels,
        "chunks": chunks,
        "batched": batched,
        "served": served,
        "full_build": full,
        "file_path": files,
        "bucket_files": bucket_files,
        "large": large,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()

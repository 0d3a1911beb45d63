"""ctypes binding of libcdbg_gpu.so (include/cdbg_gpu.h) for tests and bench.

The product is the C ABI; this module only moves numpy arrays across it.
Loading fails loudly when the HIP library has not been built: there is no
CPU fallback on the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcdbg_gpu.so")

U8P = C.POINTER(C.c_uint8)
U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "num_bases", "num_supers", "num_occ", "num_tuples", "num_nodes", "num_kmers",
        "num_color_sets", "num_frags", "frag_bytes", "cs_bits", "total_integers")] + [
        ("ms_phase1", C.c_double), ("ms_phase2", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Tuning(C.Structure):
    """Phase-2 performance knobs (cdbg_gpu_tuning); zero fields are defaults."""
    _fields_ = [("cell_target", C.c_double), ("digit_bits", C.c_uint32), ("persistent", C.c_uint32),
                ("set_slack", C.c_double), ("graph", C.c_uint32), ("gzip_device", C.c_uint32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class TuneReport(C.Structure):
    _fields_ = [("best", Tuning), ("trials", C.c_uint32), ("reserved", C.c_uint32),
                ("sample_occurrences", C.c_uint64), ("default_ms", C.c_double), ("best_ms", C.c_double)]

    def as_dict(self):
        return {"best": self.best.as_dict(), "trials": self.trials, "sample_occurrences": self.sample_occurrences,
                "default_ms": self.default_ms, "best_ms": self.best_ms}


def tuning(**kw):
    t = Tuning()
    for k, v in kw.items():
        setattr(t, k, v)
    return t


class ColorState(C.Structure):
    _fields_ = [("num_colors", C.c_uint32), ("sparse_threshold", C.c_uint32),
                ("dense_threshold", C.c_uint32), ("num_color_sets", C.c_uint64),
                ("num_bits", C.c_uint64), ("total_integers", C.c_uint64),
                ("trailing_word", C.c_uint64), ("num_kmers", C.c_uint64),
                ("num_fragments", C.c_uint64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class StitchStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("num_fragments", "open_ends", "joins", "unmatched_ends", "cycles",
                                          "num_unitigs", "unitig_kmers", "foreign_dropped")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class SynthParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("num_colors", C.c_uint32), ("genome_len", C.c_uint64),
                ("snp_rate", C.c_double), ("indel_rate", C.c_double), ("indel_max", C.c_uint32),
                ("pav_frac", C.c_double), ("pav_min", C.c_uint32), ("pav_max", C.c_uint32),
                ("n_inversions", C.c_uint32), ("inv_min", C.c_uint32), ("inv_max", C.c_uint32)]


class ServerParams(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("slots", C.c_uint32), ("keep_results", C.c_uint32),
                ("borrow_inputs", C.c_uint32)]


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("attempts", C.c_uint32), ("hbm_estimate", C.c_uint64),
                ("chunk_occurrences", C.c_uint64),
                ("submit_s", C.c_double), ("start_s", C.c_double), ("done_s", C.c_double),
                ("stats", Stats)]


# Exported symbols declared in include/cdbg_gpu.h (checked by tests/test_abi.py).
_SIGS = {
    "cdbg_gpu_last_error": (C.c_char_p, []),
    "cdbg_gpu_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "cdbg_gpu_ctx_destroy": (None, [C.c_void_p]),
    "cdbg_gpu_job_create": (C.c_int, [C.c_void_p, U8P, U64P, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint32, C.POINTER(C.c_void_p)]),
    "cdbg_gpu_job_destroy": (None, [C.c_void_p]),
    "cdbg_gpu_job_phase1": (C.c_int, [C.c_void_p]),
    "cdbg_gpu_job_phase2": (C.c_int, [C.c_void_p]),
    "cdbg_gpu_job_run": (C.c_int, [C.c_void_p]),
    "cdbg_gpu_job_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "cdbg_gpu_job_supers": (C.c_int, [C.c_void_p, C.c_int, U64P, U32P, U8P, U32P, U32P]),
    "cdbg_gpu_job_bucket_offsets": (C.c_int, [C.c_void_p, U64P]),
    "cdbg_gpu_job_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "cdbg_gpu_job_kernel_times": (C.c_int64, [C.c_void_p, C.c_char_p, C.c_uint64]),
    "cdbg_gpu_device_bytes": (C.c_uint64, []),
    "cdbg_gpu_alloc_stats": (None, [U64P]),
    "cdbg_gpu_trim_cache": (None, []),
    "cdbg_gpu_inject_oom": (None, [C.c_int64]),
    "cdbg_gpu_copy_probe": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.POINTER(C.c_double)]),
    "cdbg_gpu_job_set_chunk_tuples": (C.c_int, [C.c_void_p, C.c_uint64]),
    "cdbg_gpu_job_num_chunks": (C.c_int64, [C.c_void_p]),
    "cdbg_gpu_job_set_bucket_map": (C.c_int, [C.c_void_p, U64P, C.c_uint32]),
    "cdbg_gpu_job_set_phase1_chunk_bases": (C.c_int, [C.c_void_p, C.c_uint64]),
    "cdbg_gpu_job_num_phase1_ranges": (C.c_int64, [C.c_void_p]),
    "cdbg_gpu_phase1_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, U64P, C.c_uint32, C.c_uint32, C.c_uint32,
                                          C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p, U64P, C.c_void_p]),
    "cdbg_gpu_job_frags": (C.c_int, [C.c_void_p, U8P]),
    "cdbg_gpu_job_color_sets": (C.c_int, [C.c_void_p, U64P, U64P]),
    "cdbg_gpu_bucket_write": (C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), C.c_uint32, C.c_uint32,
                                        C.c_uint32, C.c_uint32, C.c_char_p, C.c_uint32, U64P]),
    "cdbg_gpu_bucket_process": (C.c_int, [C.c_void_p, C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                          C.c_uint32, C.c_char_p, C.c_char_p, C.POINTER(ColorState)]),
    "cdbg_gpu_bucket_process_ex": (C.c_int, [C.c_void_p, C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                             C.c_uint32, C.c_char_p, C.c_char_p, C.POINTER(ColorState),
                                             C.POINTER(C.c_uint64), C.c_uint64, C.c_int]),
    "cdbg_gpu_job_write_bucket_files": (C.c_int, [C.c_void_p, C.c_char_p, C.c_uint32, U64P]),
    "cdbg_gpu_job_from_bucket_files": (C.c_int, [C.c_void_p, C.c_char_p, C.c_uint32, C.c_uint32,
                                                 C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "cdbg_gpu_job_write_phase2_files": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p,
                                                  C.POINTER(ColorState)]),
    "cdbg_gpu_hbm_estimate": (C.c_uint64, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64]),
    "cdbg_gpu_chunk_budget": (C.c_uint64, [C.c_uint64, C.c_uint32]),
    "cdbg_gpu_server_create": (C.c_int, [C.c_void_p, C.POINTER(ServerParams), C.POINTER(C.c_void_p)]),
    "cdbg_gpu_server_submit": (C.c_int, [C.c_void_p, U8P, U64P, C.c_uint32, C.c_uint32, C.c_uint32,
                                         C.c_uint32, U64P]),
    "cdbg_gpu_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.POINTER(JobStatus)]),
    "cdbg_gpu_server_result": (C.c_int, [C.c_void_p, C.c_uint64, U8P, U64P, U64P]),
    "cdbg_gpu_server_release": (C.c_int, [C.c_void_p, C.c_uint64]),
    "cdbg_gpu_server_destroy": (None, [C.c_void_p]),
    "cdbg_stitch": (C.c_void_p, [U8P, C.c_uint64, C.c_uint32, C.c_uint32]),
    "cdbg_stitch_file": (C.c_void_p, [C.c_char_p, C.c_uint32, C.c_uint32]),
    "cdbg_gpu_stitch": (C.c_void_p, [C.c_void_p, U8P, C.c_uint64, C.c_uint32, C.c_uint32]),
    "cdbg_gpu_stitch_file": (C.c_void_p, [C.c_void_p, C.c_char_p, C.c_uint32, C.c_uint32]),
    "cdbg_frag_offsets": (C.c_int, [U8P, C.c_uint64, C.c_uint32, U64P, C.c_uint64, U64P]),
    "cdbg_gpu_job_stitch": (C.c_void_p, [C.c_void_p, C.c_uint32]),
    "cdbg_unitigs_stats": (C.c_int, [C.c_void_p, C.POINTER(StitchStats)]),
    "cdbg_unitigs_bases": (C.c_uint64, [C.c_void_p]),
    "cdbg_unitigs_export": (C.c_int, [C.c_void_p, C.c_char_p, U64P, U64P]),
    "cdbg_unitigs_free": (None, [C.c_void_p]),
    "cdbg_emit_fasta": (C.c_int, [C.c_void_p, C.c_char_p]),
    "cdbg_finalize_color_sets": (C.c_int, [C.c_char_p, C.POINTER(ColorState)]),
    "cdbg_write_metadata": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64]),
    "cdbg_load_fasta": (C.c_void_p, [C.POINTER(C.c_char_p), C.c_uint32, C.c_uint32]),
    "cdbg_load_fasta_list": (C.c_void_p, [C.c_char_p, C.c_uint32]),
    "cdbg_seqs_total_bytes": (C.c_uint64, [C.c_void_p]),
    "cdbg_seqs_num_colors": (C.c_uint32, [C.c_void_p]),
    "cdbg_seqs_bytes": (U8P, [C.c_void_p]),
    "cdbg_seqs_offsets": (U64P, [C.c_void_p]),
    "cdbg_seqs_free": (None, [C.c_void_p]),
    "cdbg_lz4_compress": (C.c_int64, [U8P, C.c_uint64, U8P, C.c_uint64]),
    "cdbg_lz4_decompress": (C.c_int64, [U8P, C.c_uint64, U8P, C.c_uint64]),
    "cdbg_gpu_job_create_from_files": (C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), C.c_uint32, C.c_uint32,
                                               C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "cdbg_gpu_job_input": (C.c_int64, [C.c_void_p, U8P, C.c_uint64, U64P]),
    "cdbg_gpu_job_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "cdbg_gpu_job_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "cdbg_gpu_ctx_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "cdbg_gpu_job_autotune": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(TuneReport)]),
    "cdbg_gpu_lz4_compress_frames": (C.c_int, [C.c_void_p, U8P, U64P, C.c_uint32, U8P, C.c_uint64, U64P]),
    "cdbg_gpu_lz4_decompress_frames": (C.c_int, [C.c_void_p, U8P, U64P, C.c_uint32, U8P, U64P]),
    "cdbg_synth_create": (C.c_void_p, [C.POINTER(SynthParams), C.c_uint32]),
    "cdbg_synth_total_bytes": (C.c_uint64, [C.c_void_p]),
    "cdbg_synth_bytes": (U8P, [C.c_void_p]),
    "cdbg_synth_offsets": (U64P, [C.c_void_p]),
    "cdbg_synth_free": (None, [C.c_void_p]),
    "cdbg_synth_write_fasta": (C.c_int, [C.c_void_p, C.c_uint32, C.c_char_p, C.c_int]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().cdbg_gpu_last_error().decode())


def synth(seed, num_colors, genome_len, snp_rate=0.01, indel_rate=0.001, indel_max=10,
          pav_frac=0.0, pav_min=1000, pav_max=5000, n_inversions=0, inv_min=10000,
          inv_max=100000, threads=8):
    """Generate a synthetic pangenome; returns (bytes u8 array, color_off u64 array)."""
    p = SynthParams(seed, num_colors, genome_len, snp_rate, indel_rate, indel_max, pav_frac,
                    pav_min, pav_max, n_inversions, inv_min, inv_max)
    h = lib().cdbg_synth_create(C.byref(p), threads)
    if not h:
        raise RuntimeError(lib().cdbg_gpu_last_error().decode())
    try:
        n = lib().cdbg_synth_total_bytes(h)
        data = np.ctypeslib.as_array(lib().cdbg_synth_bytes(h), shape=(n,)).copy()
        off = np.ctypeslib.as_array(lib().cdbg_synth_offsets(h), shape=(num_colors + 1,)).copy()
    finally:
        lib().cdbg_synth_free(h)
    return data, off


def synth_fasta_files(d, num_colors, genome_len, gz=True, threads=8, seed=1, **mut):
    """Writes the synthetic pangenome as one FASTA file per colour (80-column
    lines, gzip level 6 with gz) under directory d; returns the paths."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(d, exist_ok=True)
    p = SynthParams(seed, num_colors, genome_len, mut.get("snp_rate", 0.01), mut.get("indel_rate", 0.001),
                    mut.get("indel_max", 10), mut.get("pav_frac", 0.0), mut.get("pav_min", 1000),
                    mut.get("pav_max", 5000), mut.get("n_inversions", 0), mut.get("inv_min", 10000),
                    mut.get("inv_max", 100000))
    h = lib().cdbg_synth_create(C.byref(p), threads)
    if not h:
        raise RuntimeError(lib().cdbg_gpu_last_error().decode())
    paths = [os.path.join(d, f"color_{c}.fa" + (".gz" if gz else "")) for c in range(num_colors)]
    try:
        def one(c):
            _check(lib().cdbg_synth_write_fasta(h, c, os.fsencode(paths[c]), 1 if gz else 0))
        with ThreadPoolExecutor(max(1, threads)) as ex:
            list(ex.map(one, range(num_colors)))
    finally:
        lib().cdbg_synth_free(h)
    return paths


def _cstrs(items):
    arr = (C.c_char_p * len(items))(*[os.fsencode(x) for x in items])
    return arr


def load_fasta(files, threads=8):
    """Read FASTA/FASTQ files (kseq semantics); returns (bytes, color_off)."""
    h = lib().cdbg_load_fasta(_cstrs(files), len(files), threads)
    if not h:
        raise RuntimeError(lib().cdbg_gpu_last_error().decode())
    try:
        n = lib().cdbg_seqs_total_bytes(h)
        nc = lib().cdbg_seqs_num_colors(h)
        data = (np.ctypeslib.as_array(lib().cdbg_seqs_bytes(h), shape=(n,)).copy() if n
                else np.zeros(0, np.uint8))
        off = np.ctypeslib.as_array(lib().cdbg_seqs_offsets(h), shape=(nc + 1,)).copy()
    finally:
        lib().cdbg_seqs_free(h)
    return data, off


class Unitigs:
    """Monochromatic unitigs of a stitch (cdbg_stitch / cdbg_stitch_file), in
    emit order: ascending colour-set id."""

    def __init__(self, h):
        if not h:
            raise RuntimeError(lib().cdbg_gpu_last_error().decode())
        self.h = C.c_void_p(h)

    def close(self):
        if self.h:
            lib().cdbg_unitigs_free(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def stats(self):
        st = StitchStats()
        _check(lib().cdbg_unitigs_stats(self.h, C.byref(st)))
        return st.as_dict()

    def export(self):
        """(bases bytes, offsets u64[n+1], cids u64[n])."""
        n = self.stats()["num_unitigs"]
        nb = lib().cdbg_unitigs_bases(self.h)
        seq = C.create_string_buffer(max(nb, 1))
        offs = np.zeros(n + 1, np.uint64)
        cids = np.zeros(n, np.uint64)
        _check(lib().cdbg_unitigs_export(self.h, seq, _ptr(offs, U64P), _ptr(cids, U64P)))
        return seq.raw[:nb], offs, cids

    def unitigs(self):
        """List of (sequence bytes, cid)."""
        seq, offs, cids = self.export()
        return [(seq[int(offs[i]):int(offs[i + 1])], int(cids[i])) for i in range(len(cids))]

    def emit_fasta(self, out_basename):
        _check(lib().cdbg_emit_fasta(self.h, os.fsencode(out_basename)))


def stitch(frags, k, threads=8):
    """Phase 3 over a frag_unitigs.bin byte stream (numpy u8 array or bytes)."""
    a = np.ascontiguousarray(np.frombuffer(bytes(frags), np.uint8) if isinstance(frags, (bytes, bytearray))
                             else frags, dtype=np.uint8)
    return Unitigs(lib().cdbg_stitch(_ptr(a, U8P), len(a), k, threads))


def stitch_file(path, k, threads=8):
    return Unitigs(lib().cdbg_stitch_file(os.fsencode(path), k, threads))


def _u8(frags):
    return np.ascontiguousarray(np.frombuffer(bytes(frags), np.uint8) if isinstance(frags, (bytes, bytearray))
                                else frags, dtype=np.uint8)


def frag_offsets(frags, threads=8):
    """Start of every record of a frag_unitigs.bin byte stream (u64 array)."""
    a = _u8(frags)
    cnt = C.c_uint64()
    cap = len(a) // 9 + 1  # a record takes at least 9 bytes
    out = np.zeros(cap, np.uint64)
    _check(lib().cdbg_frag_offsets(_ptr(a, U8P), len(a), threads, _ptr(out, U64P), cap, C.byref(cnt)))
    return out[:cnt.value]


def finalize_color_sets(cs_path, state):
    """Finish <out>.color_sets from phase 2's files and colour state (a dict
    from bucket_process / write_phase2_files)."""
    st = ColorState(**{n: state[n] for n, _ in ColorState._fields_})
    _check(lib().cdbg_finalize_color_sets(os.fsencode(cs_path), C.byref(st)))


def write_metadata(out_basename, k, num_kmers, num_colors, num_unitigs, num_color_sets):
    _check(lib().cdbg_write_metadata(os.fsencode(out_basename), k, num_kmers, num_colors, num_unitigs,
                                     num_color_sets))


def lz4_compress(src):
    src = np.ascontiguousarray(np.frombuffer(bytes(src), np.uint8))
    cap = len(src) + len(src) // 255 + 64
    out = np.zeros(cap, np.uint8)
    n = lib().cdbg_lz4_compress(_ptr(src, U8P), len(src), _ptr(out, U8P), cap)
    if n < 0:
        raise RuntimeError(lib().cdbg_gpu_last_error().decode())
    return out[:n].tobytes()


def lz4_decompress(src, size):
    src = np.ascontiguousarray(np.frombuffer(bytes(src), np.uint8))
    out = np.zeros(max(size, 1), np.uint8)
    n = lib().cdbg_lz4_decompress(_ptr(src, U8P), len(src), _ptr(out, U8P), size)
    if n < 0:
        raise RuntimeError(lib().cdbg_gpu_last_error().decode())
    return out[:n].tobytes()


class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        _check(lib().cdbg_gpu_ctx_create(device, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().cdbg_gpu_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def copy_probe(self, nbytes, iters=5):
        """Measured read+write GB/s of a plain device copy of nbytes."""
        g = C.c_double()
        _check(lib().cdbg_gpu_copy_probe(self.h, nbytes, iters, C.byref(g)))
        return g.value

    def job(self, data, color_off, k, m, num_buckets):
        return Job(self, data, color_off, k, m, num_buckets)

    def job_from_files(self, files, k, m, num_buckets, threads=8):
        """A job over FASTA/FASTQ files (plain or gzip), decoded on the device
        when it can be (FASTQ goes to the host reader)."""
        j = Job.__new__(Job)
        j.num_buckets = num_buckets
        j.h = C.c_void_p()
        _check(lib().cdbg_gpu_job_create_from_files(self.h, _cstrs(files), len(files), k, m, num_buckets, threads,
                                                    C.byref(j.h)))
        return j

    def job_from_bucket_files(self, d, num_buckets, k, num_colors, threads=8):
        j = Job.__new__(Job)
        j.num_buckets = num_buckets
        j.h = C.c_void_p()
        _check(lib().cdbg_gpu_job_from_bucket_files(self.h, os.fsencode(d), num_buckets, k, num_colors,
                                                    threads, C.byref(j.h)))
        return j

    def server(self, slots=4, hbm_budget=0, keep_results=True, borrow_inputs=True):
        return Server(self, slots, hbm_budget, keep_results, borrow_inputs)

    def set_tuning(self, t=None, **kw):
        """Default tuning of jobs created on this context from now on."""
        _check(lib().cdbg_gpu_ctx_set_tuning(self.h, C.byref(t if t is not None else tuning(**kw))))

    def stitch(self, frags, k, threads=8):
        """Phase 3 on the device over a frag_unitigs.bin byte stream: the same
        unitigs, order and statistics as stitch()."""
        a = _u8(frags)
        return Unitigs(lib().cdbg_gpu_stitch(self.h, _ptr(a, U8P), len(a), k, threads))

    def stitch_file(self, path, k, threads=8):
        return Unitigs(lib().cdbg_gpu_stitch_file(self.h, os.fsencode(path), k, threads))

    def lz4_compress_frames(self, frames):
        """LZ4 blocks of the byte strings `frames` (each <= 64 KiB), compressed
        on the device, one lane per frame."""
        frames = [bytes(f) for f in frames]
        src = np.frombuffer(b"".join(frames) + bytes(8), np.uint8).copy()
        off = np.zeros(len(frames) + 1, np.uint64)
        off[1:] = np.cumsum([len(f) for f in frames], dtype=np.uint64)
        cap = int(sum(len(f) + len(f) // 255 + 16 for f in frames)) + 1
        dst = np.zeros(cap, np.uint8)
        coff = np.zeros(len(frames) + 1, np.uint64)
        _check(lib().cdbg_gpu_lz4_compress_frames(self.h, _ptr(src, U8P), _ptr(off, U64P), len(frames),
                                                  _ptr(dst, U8P), cap, _ptr(coff, U64P)))
        return [dst[int(coff[i]):int(coff[i + 1])].tobytes() for i in range(len(frames))]

    def lz4_decompress_frames(self, blocks, sizes):
        """Device decode of LZ4 blocks; block i must decode to sizes[i] bytes."""
        blocks = [bytes(b) for b in blocks]
        src = np.frombuffer(b"".join(blocks) + bytes(8), np.uint8).copy()
        coff = np.zeros(len(blocks) + 1, np.uint64)
        coff[1:] = np.cumsum([len(b) for b in blocks], dtype=np.uint64)
        doff = np.zeros(len(blocks) + 1, np.uint64)
        doff[1:] = np.cumsum(sizes, dtype=np.uint64)
        dst = np.zeros(int(doff[-1]) + 1, np.uint8)
        _check(lib().cdbg_gpu_lz4_decompress_frames(self.h, _ptr(src, U8P), _ptr(coff, U64P), len(blocks),
                                                    _ptr(dst, U8P), _ptr(doff, U64P)))
        return [dst[int(doff[i]):int(doff[i + 1])].tobytes() for i in range(len(blocks))]

    def bucket_write(self, files, k, m, num_buckets, tmp_dir, threads=8):
        unc = np.zeros(num_buckets, np.uint64)
        _check(lib().cdbg_gpu_bucket_write(self.h, _cstrs(files), len(files), k, m, num_buckets,
                                           os.fsencode(tmp_dir), threads, _ptr(unc, U64P)))
        return unc

    def bucket_process(self, tmp_dir, num_buckets, k, num_colors, frag_path, cs_path, threads=8):
        st = ColorState()
        _check(lib().cdbg_gpu_bucket_process(self.h, os.fsencode(tmp_dir), num_buckets, k, num_colors,
                                             threads, os.fsencode(frag_path), os.fsencode(cs_path),
                                             C.byref(st)))
        return st.as_dict()

    def bucket_process_ex(self, tmp_dir, num_buckets, k, num_colors, frag_path, cs_path, threads=8,
                          mem_budget_bytes=0, delete_consumed=False):
        """bucket_process with process_buckets' progress counter, memory
        budget and delete_consumed_buckets; returns (colour state, buckets done)."""
        st = ColorState()
        done = C.c_uint64(0)
        _check(lib().cdbg_gpu_bucket_process_ex(self.h, os.fsencode(tmp_dir), num_buckets, k, num_colors,
                                                threads, os.fsencode(frag_path), os.fsencode(cs_path),
                                                C.byref(st), C.byref(done), mem_budget_bytes,
                                                1 if delete_consumed else 0))
        return st.as_dict(), done.value


class Job:
    def __init__(self, ctx, data, color_off, k, m, num_buckets):
        data = np.ascontiguousarray(data, dtype=np.uint8)
        off = np.ascontiguousarray(color_off, dtype=np.uint64)
        self.num_buckets = num_buckets
        self.h = C.c_void_p()
        _check(lib().cdbg_gpu_job_create(ctx.h, _ptr(data, U8P), _ptr(off, U64P), len(off) - 1,
                                         k, m, num_buckets, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().cdbg_gpu_job_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def phase1(self):
        _check(lib().cdbg_gpu_job_phase1(self.h))

    def phase2(self):
        _check(lib().cdbg_gpu_job_phase2(self.h))

    def run(self):
        _check(lib().cdbg_gpu_job_run(self.h))

    def stats(self):
        s = Stats()
        _check(lib().cdbg_gpu_job_stats(self.h, C.byref(s)))
        return s.as_dict()

    def supers(self, which=0):
        n = self.stats()["num_supers"]
        start = np.zeros(n, np.uint64)
        ln = np.zeros(n, np.uint32)
        fl = np.zeros(n, np.uint8)
        bk = np.zeros(n, np.uint32)
        col = np.zeros(n, np.uint32)
        _check(lib().cdbg_gpu_job_supers(self.h, which, _ptr(start, U64P), _ptr(ln, U32P),
                                         _ptr(fl, U8P), _ptr(bk, U32P), _ptr(col, U32P)))
        return {"start": start, "len": ln, "flags": fl, "bucket": bk, "color": col}

    def write_bucket_files(self, d, threads=8):
        unc = np.zeros(self.num_buckets, np.uint64)
        _check(lib().cdbg_gpu_job_write_bucket_files(self.h, os.fsencode(d), threads, _ptr(unc, U64P)))
        return unc

    def stitch(self, threads=8):
        """Phase 3 on the device over this job's fragment stream where it lies
        (no frag_unitigs.bin)."""
        return Unitigs(lib().cdbg_gpu_job_stitch(self.h, threads))

    def write_phase2_files(self, frag_path, cs_path):
        st = ColorState()
        _check(lib().cdbg_gpu_job_write_phase2_files(self.h, os.fsencode(frag_path) if frag_path else None,
                                                     os.fsencode(cs_path), C.byref(st)))
        return st.as_dict()

    def set_chunk_tuples(self, n):
        _check(lib().cdbg_gpu_job_set_chunk_tuples(self.h, n))

    def set_phase1_chunk_bases(self, n):
        _check(lib().cdbg_gpu_job_set_phase1_chunk_bases(self.h, n))

    def num_phase1_ranges(self):
        return lib().cdbg_gpu_job_num_phase1_ranges(self.h)

    def num_chunks(self):
        return lib().cdbg_gpu_job_num_chunks(self.h)

    def set_bucket_map(self, cuts):
        """Phase-1 bucket map (cdbg_gpu_job_set_bucket_map); None restores
        the modulo map."""
        if cuts is None:
            _check(lib().cdbg_gpu_job_set_bucket_map(self.h, None, 0))
            return
        c = np.ascontiguousarray(cuts, dtype=np.uint64)
        _check(lib().cdbg_gpu_job_set_bucket_map(self.h, _ptr(c, U64P), len(c)))

    def input(self):
        """(bytes, colour offsets, decoded on the device?) of the job's input."""
        n = lib().cdbg_gpu_job_input(self.h, None, 0, None)
        if n < 0:
            raise RuntimeError(lib().cdbg_gpu_last_error().decode())
        on_dev = bool(n >> 62)
        n &= (1 << 62) - 1
        buf = np.zeros(max(n, 1), np.uint8)
        nc = lib().cdbg_gpu_job_input(self.h, _ptr(buf, U8P), n, None)
        if nc < 0:
            raise RuntimeError(lib().cdbg_gpu_last_error().decode())
        return buf[:n], on_dev

    def input_offsets(self, num_colors):
        off = np.zeros(num_colors + 1, np.uint64)
        if lib().cdbg_gpu_job_input(self.h, None, 0, _ptr(off, U64P)) < 0:
            raise RuntimeError(lib().cdbg_gpu_last_error().decode())
        return off

    def set_tuning(self, t=None, **kw):
        _check(lib().cdbg_gpu_job_set_tuning(self.h, C.byref(t if t is not None else tuning(**kw))))

    def get_tuning(self):
        t = Tuning()
        _check(lib().cdbg_gpu_job_get_tuning(self.h, C.byref(t)))
        return t

    def autotune(self, sample_occurrences=0):
        """Times phase 2 over a sample under candidate knobs and keeps the
        fastest (phase 1 must have run; run phase 2 again afterwards)."""
        r = TuneReport()
        _check(lib().cdbg_gpu_job_autotune(self.h, sample_occurrences, C.byref(r)))
        return r.as_dict()

    def set_profiling(self, on=True):
        _check(lib().cdbg_gpu_job_set_profiling(self.h, 1 if on else 0))

    def kernel_times(self):
        """{kernel: (launches, total_ms)} accumulated since profiling was enabled."""
        n = lib().cdbg_gpu_job_kernel_times(self.h, None, 0)
        buf = C.create_string_buffer(int(n) + 1)
        lib().cdbg_gpu_job_kernel_times(self.h, buf, n + 1)
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            out[name] = (int(cnt), float(ms))
        return out

    def bucket_offsets(self):
        out = np.zeros(self.num_buckets + 1, np.uint64)
        _check(lib().cdbg_gpu_job_bucket_offsets(self.h, _ptr(out, U64P)))
        return out

    def frags(self):
        n = self.stats()["frag_bytes"]
        out = np.zeros(n, np.uint8)
        _check(lib().cdbg_gpu_job_frags(self.h, _ptr(out, U8P)))
        return out

    def color_sets(self):
        s = self.stats()
        words = np.zeros((s["cs_bits"] + 63) // 64, np.uint64)
        offs = np.zeros(s["num_color_sets"], np.uint64)
        _check(lib().cdbg_gpu_job_color_sets(self.h, _ptr(words, U64P), _ptr(offs, U64P)))
        return words, offs


def zipf_bucket_cuts(num_buckets, s):
    """Cuts of a bucket map (cdbg_gpu_job_set_bucket_map) under which bucket b
    receives a share of the minimizers proportional to 1 / (b + 1)^s (Zipf):
    num_buckets - 1 ascending values below 2^32."""
    w = [1.0 / (b + 1) ** s for b in range(num_buckets)]
    tot = sum(w)
    cuts, acc = [], 0.0
    for b in range(num_buckets - 1):
        acc += w[b]
        cuts.append(min(int(acc / tot * 2.0 ** 32), (1 << 32) - 1))
    return np.array(cuts, dtype=np.uint64)


def alloc_stats():
    """Device block cache counters (cdbg_gpu_alloc_stats)."""
    v = np.zeros(8, np.uint64)
    lib().cdbg_gpu_alloc_stats(_ptr(v, U64P))
    keys = ("cached_bytes", "hits", "mallocs", "malloc_bytes", "malloc_ns", "trims", "trim_bytes", "trim_ns")
    return dict(zip(keys, (int(x) for x in v)))


def hbm_estimate(bases, k=31, m=21, chunk_occurrences=0):
    return lib().cdbg_gpu_hbm_estimate(bases, k, m, chunk_occurrences)


def chunk_budget(usable_bytes, k):
    """Phase-2 chunk (k-mer occurrences) the automatic budget plans for
    `usable_bytes` of free device memory at k-mer length k."""
    return lib().cdbg_gpu_chunk_budget(usable_bytes, k)


class Server:
    """Served mode (include/cdbg_gpu.h): submit jobs, poll, fetch results.

    With borrow_inputs (the default) the server reads each request's sequence
    array in place; this object keeps a reference to it until the job is done
    (or released, or the server closed), so the caller need not."""

    def __init__(self, ctx, slots=4, hbm_budget=0, keep_results=True, borrow_inputs=True):
        p = ServerParams(hbm_budget, slots, 1 if keep_results else 0, 1 if borrow_inputs else 0)
        self.h = C.c_void_p()
        self._held = {}  # job id -> sequence array the server reads in place
        _check(lib().cdbg_gpu_server_create(ctx.h, C.byref(p), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().cdbg_gpu_server_destroy(self.h)  # joins the workers: no job reads its input after this
            self.h = C.c_void_p()
        self._held.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, data, color_off, k, m, num_buckets):
        data = np.ascontiguousarray(data, dtype=np.uint8)
        off = np.ascontiguousarray(color_off, dtype=np.uint64)
        jid = C.c_uint64()
        _check(lib().cdbg_gpu_server_submit(self.h, _ptr(data, U8P), _ptr(off, U64P), len(off) - 1, k, m,
                                            num_buckets, C.byref(jid)))
        self._held[jid.value] = data
        return jid.value

    def poll(self, jid, wait=False):
        st = JobStatus()
        rc = lib().cdbg_gpu_server_poll(self.h, jid, 1 if wait else 0, C.byref(st))
        if st.state in (2, -1):  # the worker is done with the input (set before the state)
            self._held.pop(jid, None)
        _check(rc)
        return {"state": st.state, "attempts": st.attempts, "hbm_estimate": st.hbm_estimate,
                "chunk": st.chunk_occurrences,
                "submit_s": st.submit_s, "start_s": st.start_s, "done_s": st.done_s,
                "stats": st.stats.as_dict()}

    def result(self, jid):
        """(frags u8, colour-set words u64, class offsets u64) of a finished job."""
        s = self.poll(jid)["stats"]
        frags = np.zeros(s["frag_bytes"], np.uint8)
        words = np.zeros((s["cs_bits"] + 63) // 64, np.uint64)
        offs = np.zeros(s["num_color_sets"], np.uint64)
        _check(lib().cdbg_gpu_server_result(self.h, jid, _ptr(frags, U8P), _ptr(words, U64P), _ptr(offs, U64P)))
        return frags, words, offs

    def release(self, jid):
        _check(lib().cdbg_gpu_server_release(self.h, jid))
        self._held.pop(jid, None)

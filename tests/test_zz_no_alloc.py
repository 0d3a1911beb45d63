"""cdbg_gpu_phase1_noalloc (the allocation-free phase 1) through the C ABI:
its records and bucket offsets equal those of cdbg_gpu_job_phase1 on the same
input, on two calls in a row (the second with a smaller input and another k),
without a single device allocation from the library's block cache; inputs
above the limit and a too-small output capacity are rejected with a message.
Device buffers for the inputs and outputs are the test's own (hipMalloc
through the HIP runtime the library is linked against)."""
import ctypes as C

import numpy as np
import pytest

from conftest import cdbg_gpu as G

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    G.lib()  # the library's HIP runtime is loaded first; this handle is the same library
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class Dev:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(nbytes, 1)) == 0

    def put(self, a):
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, n, dtype):
        a = np.zeros(n, dtype)
        if n:
            assert self.hip.hipMemcpy(a.ctypes.data, self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(gpu_ctx, hip, data, off, k, m, B, max_supers):
    bufs = [Dev(hip, x) for x in (len(data), 8 * max_supers, 4 * max_supers, max_supers, 4 * max_supers,
                                  4 * max_supers, 8 * (B + 1))]
    try:
        bufs[0].put(np.ascontiguousarray(data))
        n = np.zeros(1, np.uint64)
        o = np.ascontiguousarray(off, np.uint64)
        rc = G.lib().cdbg_gpu_phase1_noalloc(gpu_ctx.h, bufs[0].p, G._ptr(o, G.U64P), len(o) - 1, k, m, B,
                                             max_supers, bufs[1].p, bufs[2].p, bufs[3].p, bufs[4].p, bufs[5].p,
                                             bufs[6].p, G._ptr(n, G.U64P), None)
        if rc:
            raise RuntimeError(G.lib().cdbg_gpu_last_error().decode())
        n = int(n[0])
        got = {"start": bufs[1].get(n, np.uint64), "len": bufs[2].get(n, np.uint32),
               "flags": bufs[3].get(n, np.uint8), "bucket": bufs[4].get(n, np.uint32),
               "color": bufs[5].get(n, np.uint32)}
        return got, bufs[6].get(B + 1, np.uint64)
    finally:
        for b in bufs:
            b.free()


def _job(gpu_ctx, data, off, k, m, B):
    with gpu_ctx.job(data, off, k, m, B) as j:
        j.phase1()
        return j.supers(1), j.bucket_offsets()


def test_noalloc_phase1_matches_job_phase1(gpu_ctx, hip):
    cases = [(dict(seed=31, num_colors=6, genome_len=500_000), 31, 21, 1024),  # 3 Mbp
             (dict(seed=32, num_colors=2, genome_len=60_000), 63, 27, 100)]    # then a smaller one
    for params, k, m, B in cases:
        data, off = G.synth(snp_rate=0.01, indel_rate=0.001, **params)
        want, want_off = _job(gpu_ctx, data, off, k, m, B)
        n_want = len(want["start"])
        before = G.alloc_stats()
        got, got_off = _noalloc(gpu_ctx, hip, data, off, k, m, B, n_want + 17)
        after = G.alloc_stats()
        assert after["mallocs"] == before["mallocs"] and after["hits"] == before["hits"], "the call allocated"
        assert np.array_equal(got_off, want_off)
        for key in ("start", "len", "flags", "bucket", "color"):
            assert np.array_equal(got[key], want[key]), key


def test_noalloc_phase1_tiny_input(gpu_ctx, hip):
    # a handful of records: the outputs are smaller than a buffer's 256-byte minimum
    data, off = G.synth(seed=33, num_colors=1, genome_len=300)
    want, want_off = _job(gpu_ctx, data, off, 31, 19, 4)
    got, got_off = _noalloc(gpu_ctx, hip, data, off, 31, 19, 4, len(want["start"]))
    assert np.array_equal(got_off, want_off)
    for key in want:
        assert np.array_equal(got[key], want[key]), key


def test_noalloc_phase1_rejects(gpu_ctx, hip):
    data, off = G.synth(seed=34, num_colors=2, genome_len=20_000)
    want, _ = _job(gpu_ctx, data, off, 31, 19, 16)
    with pytest.raises(RuntimeError, match="more than max_supers"):
        _noalloc(gpu_ctx, hip, data, off, 31, 19, 16, len(want["start"]) - 1)
    big = np.zeros(5 << 20, np.uint8) + ord("A")
    with pytest.raises(RuntimeError, match="allocation-free limit"):
        _noalloc(gpu_ctx, hip, big, np.array([0, len(big)], np.uint64), 31, 19, 16, 16)
    # the entry point still works after the failures
    got, _ = _noalloc(gpu_ctx, hip, data, off, 31, 19, 16, len(want["start"]))
    assert np.array_equal(got["start"], want["start"])

"""ctypes binding of the oracle (liboracle.so) and of the reference-header
checker shim (_ref/libref.so).  TEST INFRASTRUCTURE ONLY: imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

U8P = C.POINTER(C.c_uint8)
U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)


def _p(a, t):
    return a.ctypes.data_as(t)


_orc = None


def orc():
    global _orc
    if _orc is None:
        L = C.CDLL(ORACLE_SO)
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_best_m": (C.c_uint32, [C.c_uint32]),
            "orc_mmer_hashes": (None, [U8P, C.c_uint64, C.c_uint32, U64P]),
            "orc_window_min": (None, [U8P, C.c_uint64, C.c_uint32, C.c_uint32, U64P]),
            "orc_revcomp64": (C.c_uint64, [C.c_uint64, C.c_uint32]),
            "orc_revcomp128": (None, [C.c_uint64, C.c_uint64, C.c_uint32, U64P]),
            "orc_encode_set": (C.c_uint64, [U32P, C.c_uint64, C.c_uint32, U64P, C.c_uint64]),
            "orc_write_delta": (None, [C.c_uint64, U64P, U64P]),
            "orc_supers": (C.c_uint64, [U8P, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32,
                                        C.c_uint32, C.c_uint32, C.c_uint64, U64P, U32P, U8P, U32P]),
            "orc_build": (C.c_void_p, [U8P, U64P, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.c_uint32, C.c_uint32]),
            "orc_build_mapped": (C.c_void_p, [U8P, U64P, C.c_uint32, C.c_uint32, C.c_uint32,
                                              C.c_uint32, C.c_uint32, U64P, C.c_uint32]),
            "orc_process_bucket_files": (C.c_void_p, [C.c_char_p, C.c_uint32, C.c_uint32,
                                                      C.c_uint32, C.c_uint32]),
            "orc_free": (None, [C.c_void_p]),
            "orc_stat": (C.c_uint64, [C.c_void_p, C.c_int]),
            "orc_time": (C.c_double, [C.c_void_p, C.c_int]),
            "orc_copy_frags": (None, [C.c_void_p, U8P]),
            "orc_copy_cs": (None, [C.c_void_p, U64P, U64P]),
            "orc_copy_supers": (None, [C.c_void_p, C.c_int, U64P, U32P, U8P, U32P, U32P]),
            "orc_lz4_decode": (C.c_int, [U8P, C.c_uint64, U8P, C.c_uint64, U64P]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _orc = L
    return _orc


def ref_available():
    return os.path.exists(REF_SO)


_ref = None


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_SO)
        sig = {
            "ref_compute_best_m": (C.c_uint32, [C.c_uint32]),
            "ref_nuc_to_2bit": (C.c_uint8, [C.c_char]),
            "ref_rolled_hashes": (None, [U8P, C.c_uint64, C.c_uint32, U64P]),
            "ref_window_min": (None, [U8P, C.c_uint64, C.c_uint32, C.c_uint32, U64P]),
            "ref_revcomp64": (C.c_uint64, [C.c_uint64, C.c_uint32]),
            "ref_revcomp128": (None, [C.c_uint64, C.c_uint64, C.c_uint32, U64P]),
            "ref_varint": (C.c_uint32, [C.c_uint64, U8P]),
            "ref_record": (C.c_uint32, [C.c_uint8, U32P, C.c_uint32, U8P, C.c_uint32, U8P]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _ref = L
    return _ref


STAT = {"num_kmers": 0, "num_nodes": 1, "num_occ": 2, "num_frags": 3, "num_sets": 4,
        "cs_bits": 5, "total_integers": 6, "num_supers": 7, "frag_bytes": 8}


class Build:
    """Oracle phases 1+2 over the per-occurrence bucket stream."""

    def __init__(self, data, color_off, k, m, B, threads=1, cuts=None):
        data = np.ascontiguousarray(data, np.uint8)
        off = np.ascontiguousarray(color_off, np.uint64)
        if cuts is None:
            self.h = orc().orc_build(_p(data, U8P), _p(off, U64P), len(off) - 1, k, m, B, threads)
        else:  # bucket map: B - 1 ascending cuts of (minimizer >> 1)
            c = np.ascontiguousarray(cuts, np.uint64)
            self.h = orc().orc_build_mapped(_p(data, U8P), _p(off, U64P), len(off) - 1, k, m, B, threads,
                                            _p(c, U64P), len(c))
        if not self.h:
            raise RuntimeError(orc().orc_last_error().decode())

    @classmethod
    def from_bucket_files(cls, d, B, k, num_colors, threads=1):
        self = cls.__new__(cls)
        self.h = orc().orc_process_bucket_files(d.encode(), B, k, num_colors, threads)
        if not self.h:
            raise RuntimeError(orc().orc_last_error().decode())
        return self

    def __del__(self):
        if getattr(self, "h", None):
            orc().orc_free(self.h)
            self.h = None

    def stat(self, name):
        return orc().orc_stat(self.h, STAT[name])

    def times(self):
        return orc().orc_time(self.h, 0), orc().orc_time(self.h, 1)

    def supers(self, which=0):
        n = self.stat("num_supers")
        start = np.zeros(n, np.uint64)
        ln = np.zeros(n, np.uint32)
        fl = np.zeros(n, np.uint8)
        bk = np.zeros(n, np.uint32)
        col = np.zeros(n, np.uint32)
        orc().orc_copy_supers(self.h, which, _p(start, U64P), _p(ln, U32P), _p(fl, U8P),
                              _p(bk, U32P), _p(col, U32P))
        return {"start": start, "len": ln, "flags": fl, "bucket": bk, "color": col}

    def frags(self):
        out = np.zeros(self.stat("frag_bytes"), np.uint8)
        orc().orc_copy_frags(self.h, _p(out, U8P))
        return out

    def color_sets(self):
        words = np.zeros((self.stat("cs_bits") + 63) // 64, np.uint64)
        offs = np.zeros(self.stat("num_sets"), np.uint64)
        orc().orc_copy_cs(self.h, _p(words, U64P), _p(offs, U64P))
        return words, offs


def codes_of(seq_bytes):
    """ASCII -> 2-bit codes (0xff for non-ACGT), the nuc_to_2bit mapping."""
    t = np.full(256, 0xFF, np.uint8)
    for i, c in enumerate(b"ACGT"):
        t[c] = i
        t[c + 32] = i
    return t[np.frombuffer(bytes(seq_bytes), np.uint8)]


def mmer_hashes(codes, m):
    codes = np.ascontiguousarray(codes, np.uint8)
    out = np.zeros(max(len(codes) - m + 1, 0), np.uint64)
    orc().orc_mmer_hashes(_p(codes, U8P), len(codes), m, _p(out, U64P))
    return out


def window_min(codes, k, m):
    codes = np.ascontiguousarray(codes, np.uint8)
    out = np.zeros(max(len(codes) - k + 2, 0), np.uint64)
    orc().orc_window_min(_p(codes, U8P), len(codes), k, m, _p(out, U64P))
    return out


def supers(seq, k, m, B, color=0, abs0=0):
    seq = np.ascontiguousarray(np.frombuffer(bytes(seq), np.uint8))
    cap = len(seq) + 1
    start = np.zeros(cap, np.uint64)
    ln = np.zeros(cap, np.uint32)
    fl = np.zeros(cap, np.uint8)
    bk = np.zeros(cap, np.uint32)
    n = orc().orc_supers(_p(seq, U8P), len(seq), abs0, k, m, B, color, cap, _p(start, U64P),
                         _p(ln, U32P), _p(fl, U8P), _p(bk, U32P))
    return {"start": start[:n], "len": ln[:n], "flags": fl[:n], "bucket": bk[:n]}


def encode_set(colors, num_colors):
    c = np.ascontiguousarray(colors, np.uint32)
    cap = 4 + (num_colors + 64 * 4 * (len(c) + 1)) // 64 + 8
    out = np.zeros(cap, np.uint64)
    nbits = orc().orc_encode_set(_p(c, U32P), len(c), num_colors, _p(out, U64P), cap)
    return out[: (nbits + 63) // 64], nbits

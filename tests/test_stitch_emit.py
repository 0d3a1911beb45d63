"""Phase 3 (stitch) and phase 4 (emit) on the host.

Stitch (cdbg_stitch, stitch_extmem.hpp:112-355 semantics) over the oracle's
frag_unitigs.bin:
* gives exactly the ground-truth coloured compacted dBG (test_data/verify.py:
  103-220) on mutated genomes, repeats, cycles, hairpins and branches, and
  under a skewed bucket map -- compared as k-mer sets, so orientation and the
  rotation of cycles do not matter, each unitig with its ground-truth colours;
* equals the independent test-side stitch (ccdbg_check) on a larger pangenome;
* the salmonella_10 known answer: 86,630 unitigs (runs where the reference's
  test data is present).
Emit: <out>.fa holds every unitig once in ascending colour-set id; <out>.u2c
has one bit per unitig and one set bit per colour set (emit.hpp:27-82,131-166);
the finished <out>.color_sets carries the header totals, every word of the
stream and an Elias-Fano tail that decodes (with a reader written from the
layout the writer states) to the class offsets plus the sentinel
(streaming_color_set_dict.hpp:221-294); the metadata file has the five
fields of builder.hpp:752-764.  The Elias-Fano layout follows jermp/bits,
which is absent: parity unpinned beyond this round trip.
"""
import collections
import os
import struct

import numpy as np
import pytest

import ccdbg_check as C
import oracle as O
from conftest import cdbg_gpu as G
from test_oracle_phase2 import SAL, _job, _mutated, _records


def _sets(ob, nc):
    words, offs = ob.color_sets()
    return C.decode_color_sets(words, list(offs) + [ob.stat("cs_bits")], nc)


def _cpp(ob, k):
    with G.stitch(ob.frags(), k, threads=4) as U:
        return U.unitigs(), U.stats()


def _check_truth(data, off, k, m, B, cuts=None):
    ob = O.Build(data, off, k, m, B, cuts=cuts)
    nc = len(off) - 1
    S = C.color_map(_records(data, off), k)
    masks = [sum(1 << c for c in cs) for cs in _sets(ob, nc)]
    got, st = _cpp(ob, k)
    assert st["num_unitigs"] == len(got)
    assert st["unitig_kmers"] == ob.stat("num_kmers") == len(S)
    cids = [c for _, c in got]
    assert cids == sorted(cids), "emit order is ascending cid"
    have = {}
    for seq, cid in got:
        assert C.check_path(seq, k, S)
        key = C.unitig_key(seq, k)
        assert key not in have, "unitig emitted twice"
        have[key] = masks[cid]
    want = {C.unitig_key(seq, k): mask for seq, mask, _ in C.truth_unitigs(S, k)}
    assert have == want
    return st


@pytest.mark.parametrize("k,m,B,seed", [(15, 7, 4, 1), (15, 9, 16, 2), (21, 11, 8, 3), (31, 19, 64, 4),
                                        (32, 19, 16, 5), (33, 21, 8, 6)])
def test_stitch_equals_ground_truth_mutated_genomes(k, m, B, seed):
    base = bytes(np.frombuffer(b"ACGT", np.uint8)[np.random.default_rng(seed).integers(0, 4, 3000)])
    data, off = _job([[g] for g in _mutated(base, 4, 0.02, 0.003, seed + 100)])
    st = _check_truth(data, off, k, m, B)
    assert st["joins"] > 0 and st["num_unitigs"] > 10


@pytest.mark.parametrize("k,m,B", [(11, 5, 4), (15, 7, 8), (12, 5, 3)])
def test_stitch_equals_ground_truth_repeats_cycles_and_branches(k, m, B):
    r = np.random.default_rng(5)
    unit = bytes(np.frombuffer(b"ACGT", np.uint8)[r.integers(0, 4, 40)])
    tandem = unit * 6
    hairpin = unit + C.rc(unit)
    shared = bytes(np.frombuffer(b"ACGT", np.uint8)[r.integers(0, 4, 300)])
    branchy = shared[:150] + b"A" + shared[150:] + b"NN" + shared[:150] + b"C" + shared[150:]
    ring = bytes(np.frombuffer(b"ACGT", np.uint8)[r.integers(0, 4, 200)])
    colours = [
        [tandem, b"acgtNacgtacgtacgtacgtacgt" + shared, ring + ring[:k - 1]],  # a cycle spanning buckets
        [hairpin, branchy, ring[:120]],                                        # part of the ring in two colours
        [unit * 2 + shared, b"GGGGGGGGGGGGGGGGGGGGGG"],
    ]
    data, off = _job(colours)
    _check_truth(data, off, k, m, B)


def test_stitch_equals_ground_truth_synthetic_pangenome_and_skew():
    data, off = G.synth(seed=9, num_colors=5, genome_len=4000, snp_rate=0.01, indel_rate=0.002,
                        pav_frac=0.05, pav_min=100, pav_max=400, n_inversions=1, inv_min=300, inv_max=800)
    _check_truth(data, off, 21, 9, 16)
    _check_truth(data, off, 31, 19, 256, cuts=G.zipf_bucket_cuts(256, 1.5))


def test_stitch_equals_test_side_stitch():
    data, off = G.synth(seed=19, num_colors=10, genome_len=30000, snp_rate=0.01, indel_rate=0.001)
    k = 31
    ob = O.Build(data, off, k, 19, 64, threads=4)
    got, st = _cpp(ob, k)
    py = C.split_monochromatic(C.stitch(C.parse_frags(ob.frags()), k), k)
    assert collections.Counter((C.unitig_key(s, k), c) for s, c in got) == \
        collections.Counter((C.unitig_key(s, k), c) for s, c, _ in py)
    assert st["num_fragments"] == ob.stat("num_frags")


def test_emit_is_independent_of_the_thread_count(tmp_path):
    data, off = G.synth(seed=19, num_colors=10, genome_len=30000, snp_rate=0.01, indel_rate=0.001)
    f = O.Build(data, off, 31, 19, 64, threads=4).frags()
    outs = []
    for th in (1, 3, 16):
        with G.stitch(f, 31, threads=th) as U:
            U.emit_fasta(str(tmp_path / f"e{th}"))
        outs.append(((tmp_path / f"e{th}.fa").read_bytes(), (tmp_path / f"e{th}.u2c").read_bytes()))
    assert outs[0] == outs[1] == outs[2] and len(outs[0][0]) > 100000


def test_stitch_rejects_truncated_input():
    data, off = G.synth(seed=2, num_colors=2, genome_len=3000)
    f = O.Build(data, off, 21, 11, 8).frags()
    with pytest.raises(RuntimeError, match="truncated"):
        G.stitch(f[:-3], 21)


def _phase2_files(ob, d, nc):
    """The files phase 2 leaves: <out>.color_sets = 28 zero bytes + complete
    words, its .tmp_offsets sidecar, and the colour state."""
    words, offs = ob.color_sets()
    bits = ob.stat("cs_bits")
    full = bits // 64
    cs = os.path.join(d, "out.color_sets")
    open(cs, "wb").write(bytes(28) + words[:full].tobytes())
    open(cs + ".tmp_offsets", "wb").write(offs.tobytes())
    state = dict(num_colors=nc, sparse_threshold=nc // 4, dense_threshold=3 * nc // 4,
                 num_color_sets=ob.stat("num_sets"), num_bits=bits, total_integers=ob.stat("total_integers"),
                 trailing_word=int(words[full]) if bits % 64 else 0, num_kmers=ob.stat("num_kmers"),
                 num_fragments=ob.stat("num_frags"))
    return cs, state, words, offs


def read_elias_fano(buf, p):
    """Decodes the Elias-Fano tail at buf[p:]: returns (values, end)."""
    def u64():
        nonlocal p
        v, = struct.unpack_from("<Q", buf, p)
        p += 8
        return v

    def vec(fmt, size):
        nonlocal p
        n = u64()
        v = list(struct.unpack_from("<%d%s" % (n, fmt), buf, p))
        p += n * size
        return v

    back = u64()
    hb_bits = u64()
    high = vec("Q", 8)
    d1 = (u64(), vec("q", 8), vec("H", 2), vec("Q", 8))
    d0 = (u64(), vec("q", 8), vec("H", 2), vec("Q", 8))
    n, l, mask = u64(), u64(), u64()
    low = vec("Q", 8)
    assert mask == (1 << l) - 1 and len(high) == (hb_bits + 63) // 64 and len(low) == (n * l + 63) // 64 + 1
    ones = [i for i in range(hb_bits) if (high[i >> 6] >> (i & 63)) & 1]
    assert len(ones) == n == d1[0] and d0 == (0, [], [], [])
    # the select index: one inventory entry per 1024 ones, a subblock entry per 32
    assert len(d1[1]) == (n + 1023) // 1024 and len(d1[2]) == sum((min(1024, n - b) + 31) // 32
                                                                   for b in range(0, n, 1024))
    for b, e in enumerate(d1[1]):
        if e >= 0:
            assert e == ones[1024 * b]
    lowbits = 0
    for i, w in enumerate(low):
        lowbits |= w << (64 * i)
    vals = [((ones[i] - i) << l) | ((lowbits >> (i * l)) & mask) for i in range(n)]
    assert not vals or back == vals[-1]
    return vals, p


@pytest.mark.parametrize("nc,glen", [(3, 3000), (12, 8000), (40, 2000)])
def test_emit_files(tmp_path, nc, glen):
    data, off = G.synth(seed=nc, num_colors=nc, genome_len=glen, snp_rate=0.02, indel_rate=0.002)
    k = 21
    ob = O.Build(data, off, k, 11, 32)
    cs, state, words, offs = _phase2_files(ob, str(tmp_path), nc)
    out = str(tmp_path / "out")
    with G.stitch(ob.frags(), k) as U:
        st = U.stats()
        unitigs = U.unitigs()
        U.emit_fasta(out)
    G.finalize_color_sets(cs, state)
    G.write_metadata(out, k, ob.stat("num_kmers"), nc, st["num_unitigs"], ob.stat("num_sets"))
    assert not os.path.exists(cs + ".tmp_offsets")
    # .fa
    lines = open(out + ".fa", "rb").read().split(b"\n")
    assert lines[-1] == b"" and len(lines) == 2 * len(unitigs) + 1
    fa = [(lines[2 * i + 1], int(lines[2 * i][1:])) for i in range(len(unitigs))]
    assert fa == unitigs
    # .u2c
    raw = open(out + ".u2c", "rb").read()
    nbits, nwords = struct.unpack_from("<QQ", raw)
    assert nbits == len(unitigs) and nwords == (nbits + 63) // 64 and len(raw) == 16 + 8 * nwords
    u2c = np.frombuffer(raw[16:], np.uint64)
    bits = [(int(u2c[i >> 6]) >> (i & 63)) & 1 for i in range(nbits)]
    assert sum(bits) == ob.stat("num_sets")
    cids = [c for _, c in unitigs]
    assert all(bits[i] == (i + 1 == nbits or cids[i + 1] != cids[i]) for i in range(nbits))
    assert sorted(set(cids)) == list(range(ob.stat("num_sets")))
    # .color_sets
    buf = open(cs, "rb").read()
    h = struct.unpack_from("<IIIQQ", buf)
    nw = (state["num_bits"] + 63) // 64
    assert h == (nc, nc // 4, 3 * nc // 4, state["num_bits"], nw)
    assert np.array_equal(np.frombuffer(buf[28:28 + 8 * nw], np.uint64), words[:nw])
    vals, end = read_elias_fano(buf, 28 + 8 * nw)
    assert end == len(buf)
    assert vals == [int(x) for x in offs] + [state["num_bits"]]
    # metadata
    meta = dict(x.split("=") for x in open(out + ".metadata.txt").read().split())
    assert meta == {"k": str(k), "num_kmers": str(ob.stat("num_kmers")), "num_colors": str(nc),
                    "num_unitigs": str(st["num_unitigs"]), "num_color_sets": str(ob.stat("num_sets"))}


def test_elias_fano_blocks_and_overflow(tmp_path):
    """Offsets that spread beyond 2^16 bit positions per 1024 ones take the
    overflow inventory; dense ones take the subblock inventory."""
    rng = np.random.default_rng(3)
    for gaps in (rng.integers(0, 4, 5000), rng.integers(0, 400, 3000), np.zeros(2100, np.int64)):
        offs = np.cumsum(gaps).astype(np.uint64)
        total = int(offs[-1]) + 17
        cs = str(tmp_path / "x.color_sets")
        full = total // 64
        open(cs, "wb").write(bytes(28) + bytes(8 * full))
        open(cs + ".tmp_offsets", "wb").write(offs.tobytes())
        G.finalize_color_sets(cs, dict(num_colors=9, sparse_threshold=2, dense_threshold=6,
                                       num_color_sets=len(offs), num_bits=total, total_integers=0,
                                       trailing_word=5, num_kmers=0, num_fragments=0))
        buf = open(cs, "rb").read()
        nw = (total + 63) // 64
        assert struct.unpack_from("<Q", buf, 28 + 8 * full)[0] == 5
        vals, end = read_elias_fano(buf, 28 + 8 * nw)
        assert end == len(buf) and vals == [int(x) for x in offs] + [total]


@pytest.mark.skipif(len(SAL) != 10, reason="reference test data not present")
def test_salmonella_10_unitigs():
    data, off = G.load_fasta(SAL, threads=8)
    ob = O.Build(data, off, 31, 12, 1024, threads=8)
    _, st = _cpp(ob, 31)
    assert st["num_unitigs"] == 86630
    assert st["unitig_kmers"] == ob.stat("num_kmers")

"""Pins the oracle's phase 2 (bucket_walker.hpp load/resolve/walk, the global
colour-set dictionary and the fragment records) by what the final graph must
be, not by its own formulas:

* ground truth: the coloured compacted dBG by definition
  (test_data/verify.py:103-220) over the input's canonical k-mers;
* the oracle's frag_unitigs.bin, joined at open ends and split at colour-run
  boundaries by an independent test-side stitch, must give exactly the
  ground-truth unitigs (as k-mer sets, so orientation and the rotation of
  cycles do not matter), each with its ground-truth colour set, decoded from
  the oracle's .color_sets stream by a reader written from stream.hpp:196-237;
* the dictionary holds exactly the input's distinct colour sets, once each;
* known answers on the reference's salmonella_10 data (algorithm.md:36-39,
  test_data/README.md): 86,630 unitigs, 171 colour sets, ~6.9 M k-mers
  (runs only where the reference's test data is present).
"""
import glob
import os

import numpy as np
import pytest

import oracle as O
from conftest import cdbg_gpu as G
import ccdbg_check as C

rng = np.random.default_rng(17)


def _records(data, off):
    out = []
    for c in range(len(off) - 1):
        blob = bytes(data[off[c]:off[c + 1]])
        out.append([r for r in blob.split(b"\n")])
    return out


def _final_unitigs(ob, k, num_colors):
    frags = C.parse_frags(ob.frags())
    assert len(frags) == ob.stat("num_frags")
    words, offs = ob.color_sets()  # start bit of each class
    sets = C.decode_color_sets(words, list(offs) + [ob.stat("cs_bits")], num_colors)
    assert len(sets) == ob.stat("num_sets")
    masks = [sum(1 << c for c in cs) for cs in sets]
    assert len(set(masks)) == len(masks), "a colour set is stored twice"
    pieces = C.split_monochromatic(C.stitch(frags, k), k)
    return [(seq, masks[cid]) for seq, cid, _ in pieces], masks


def _check_against_truth(data, off, k, m, B, cuts=None):
    ob = O.Build(data, off, k, m, B, cuts=cuts)
    nc = len(off) - 1
    S = C.color_map(_records(data, off), k)
    assert ob.stat("num_kmers") == len(S)
    got, masks = _final_unitigs(ob, k, nc)
    assert set(masks) == set(S.values())
    truth = C.truth_unitigs(S, k)
    want = {}
    for seq, mask, _ in truth:
        want[C.unitig_key(seq, k)] = mask
    have = {}
    for seq, mask in got:
        assert C.check_path(seq, k, S)
        key = C.unitig_key(seq, k)
        assert key not in have, "unitig emitted twice"
        have[key] = mask
    assert len(have) == len(want)
    assert have == want
    return len(want)


def _mutated(base, n, snp, indel, seed):
    r = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        s = bytearray(base)
        for i in np.nonzero(r.random(len(s)) < snp)[0]:
            s[i] = b"ACGT"[(b"ACGT".index(s[i]) + 1 + int(r.integers(0, 3))) % 4]
        for i in sorted(np.nonzero(r.random(len(s)) < indel)[0], reverse=True):
            if r.random() < 0.5:
                s[i:i + 1] = b""
            else:
                s[i:i] = bytes(np.frombuffer(b"ACGT", np.uint8)[r.integers(0, 4, int(r.integers(1, 6)))])
        out.append(bytes(s))
    return out


def _job(colour_recs):
    data, off = bytearray(), [0]
    for recs in colour_recs:
        for r in recs:
            data += r + b"\n"
        off.append(len(data))
    return np.frombuffer(bytes(data), np.uint8), np.array(off, np.uint64)


@pytest.mark.parametrize("k,m,B,seed", [(15, 7, 4, 1), (15, 9, 16, 2), (21, 11, 8, 3), (31, 19, 64, 4)])
def test_unitigs_equal_ground_truth_mutated_genomes(k, m, B, seed):
    base = bytes(np.frombuffer(b"ACGT", np.uint8)[np.random.default_rng(seed).integers(0, 4, 3000)])
    genomes = _mutated(base, 4, 0.02, 0.003, seed + 100)
    data, off = _job([[g] for g in genomes])
    assert _check_against_truth(data, off, k, m, B) > 10


@pytest.mark.parametrize("k,m,B", [(11, 5, 4), (15, 7, 8)])
def test_unitigs_equal_ground_truth_repeats_cycles_and_branches(k, m, B):
    r = np.random.default_rng(5)
    unit = bytes(np.frombuffer(b"ACGT", np.uint8)[r.integers(0, 4, 40)])
    tandem = unit * 6                                   # a cycle of 40 k-mers
    hairpin = unit + C.rc(unit)                          # self-reverse-complementary join
    shared = bytes(np.frombuffer(b"ACGT", np.uint8)[r.integers(0, 4, 300)])
    branchy = shared[:150] + b"A" + shared[150:] + b"NN" + shared[:150] + b"C" + shared[150:]
    colours = [
        [tandem, b"acgtNacgtacgtacgtacgtacgt" + shared],
        [hairpin, branchy],
        [unit * 2 + shared, b"GGGGGGGGGGGGGGGGGGGGGG"],
    ]
    data, off = _job(colours)
    _check_against_truth(data, off, k, m, B)


def test_unitigs_equal_ground_truth_synthetic_pangenome():
    data, off = G.synth(seed=9, num_colors=5, genome_len=4000, snp_rate=0.01, indel_rate=0.002,
                        pav_frac=0.05, pav_min=100, pav_max=400, n_inversions=1, inv_min=300, inv_max=800)
    _check_against_truth(data, off, 21, 9, 16)


@pytest.mark.parametrize("k,m,B,s", [(21, 9, 64, 1.1), (31, 19, 256, 1.5)])
def test_unitigs_equal_ground_truth_under_skewed_bucket_map(k, m, B, s):
    """A Zipf bucket map (every minimizer keeps one bucket) gives skewed
    bucket sizes and the same graph."""
    data, off = G.synth(seed=4, num_colors=6, genome_len=6000, snp_rate=0.01, indel_rate=0.001)
    cuts = G.zipf_bucket_cuts(B, s)
    sizes = np.bincount(O.Build(data, off, k, m, B, cuts=cuts).supers()["bucket"], minlength=B)
    plain = np.bincount(O.Build(data, off, k, m, B).supers()["bucket"], minlength=B)
    assert sizes.sum() == plain.sum()
    assert sizes[0] > 5 * plain.max()  # bucket 0 takes the largest share of minimizers
    assert _check_against_truth(data, off, k, m, B, cuts=cuts) > 10


def test_zipf_bucket_cuts_shape():
    c = G.zipf_bucket_cuts(4096, 1.1).astype(object)
    assert len(c) == 4095 and all(a <= b for a, b in zip(c, c[1:])) and c[-1] < 2 ** 32
    share = [c[0]] + [b - a for a, b in zip(c, c[1:])] + [2 ** 32 - c[-1]]
    assert 9000 < share[0] / share[-1] < 9800  # 4096^1.1


SAL = sorted(glob.glob("/root/reference/test_data/salmonella_10/*.fasta.gz"))


@pytest.mark.skipif(len(SAL) != 10, reason="reference test data not present")
def test_salmonella_10_known_answers():
    data, off = G.load_fasta(SAL, threads=8)
    k = 31
    ob = O.Build(data, off, k, 12, 1024, threads=8)  # m = compute_best_m(31)
    assert ob.stat("num_sets") == 171
    assert 6.8e6 < ob.stat("num_kmers") < 7.0e6
    got, masks = _final_unitigs(ob, k, 10)
    assert len(got) == 86630

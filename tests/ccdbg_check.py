"""Test-side reconstruction of the final coloured compacted de Bruijn graph
from phase-2 artifacts, and the ground-truth ccdBG it is compared with.

* parse_frags: frag_unitigs.bin records
  `[u8 open][u32 nruns]{[u64 cid][u32 n]}[u32 seq_len][seq]`
  (unitig_spill.hpp:445-555, parser :646-665).
* decode_color_sets: the hybrid colour-set stream, decoded the way the
  reference's reader does (stream.hpp:196-237); Elias-delta reads follow the
  LSB-first layout of jermp/bits write_delta (bit order parity unpinned).
* stitch: joins open fragment ends that share their canonical boundary k-mer
  and splits the joined tigs at colour-run boundaries, dropping foreign
  placeholders at extremities (stitch_extmem.hpp:112-228 semantics).  A plain
  chain walk, not the reference's round-based external-memory join.
* truth_unitigs: the ccdBG by definition, as test_data/verify.py:103-220
  states it: maximal non-branching, monochromatic paths over the canonical
  k-mer set, degrees counted over distinct canonical neighbours.
"""
import struct

FOREIGN = (1 << 64) - 1
OPEN_LEFT, OPEN_RIGHT = 1, 2
_RC = bytes.maketrans(b"ACGT", b"TGCA")


def rc(s):
    return s.translate(_RC)[::-1]


def canon(s):
    r = rc(s)
    return s if s <= r else r


def parse_frags(buf):
    buf = bytes(buf)
    out, p = [], 0
    while p < len(buf):
        of, nr = struct.unpack_from("<BI", buf, p)
        p += 5
        runs = []
        for _ in range(nr):
            cid, n = struct.unpack_from("<QI", buf, p)
            p += 12
            runs.append((cid, n))
        ln, = struct.unpack_from("<I", buf, p)
        p += 4
        out.append((of, runs, buf[p:p + ln]))
        p += ln
    return out


class _Bits:
    def __init__(self, words):
        self.w = [int(x) for x in words]
        self.pos = 0

    def bit(self):
        b = (self.w[self.pos >> 6] >> (self.pos & 63)) & 1
        self.pos += 1
        return b

    def read(self, n):
        v = 0
        for i in range(n):
            v |= self.bit() << i
        return v

    def gamma(self):
        z = 0
        while self.bit() == 0:
            z += 1
        return ((1 << z) | self.read(z)) - 1

    def delta(self):
        b = self.gamma()
        return ((1 << b) | self.read(b)) - 1


def decode_color_sets(words, offsets, num_colors):
    """List of colour lists, one per class id (stream.hpp:196-237)."""
    sparse, dense = num_colors // 4, (3 * num_colors) // 4
    bs = _Bits(list(words) + [0])
    sets = []
    for i in range(len(offsets) - 1):
        bs.pos = int(offsets[i])
        n = bs.delta()
        cs = []
        if n < sparse:
            prev = bs.delta()
            cs.append(prev)
            for _ in range(1, n):
                prev = bs.delta() + prev + 1
                cs.append(prev)
        elif n < dense:
            cs = [c for c in range(num_colors) if bs.bit()]
        else:
            curr = nxt = 0
            for _ in range(num_colors - n):
                nxt += bs.delta()
                cs.extend(range(curr, nxt))
                curr = nxt + 1
                nxt += 1
            cs.extend(range(curr, num_colors))
        assert len(cs) == n and bs.pos == int(offsets[i + 1]), f"class {i}"
        sets.append(cs)
    return sets


def _oriented(frag, reverse):
    of, runs, seq = frag
    cids = [c for c, n in runs for _ in range(n)]
    if not reverse:
        return seq, cids, of
    of2 = ((of & OPEN_LEFT) << 1) | ((of & OPEN_RIGHT) >> 1)
    return rc(seq), cids[::-1], of2


def stitch(frags, k):
    """Joined tigs as (sequence, per-k-mer cid list, circular)."""
    ends = {}
    for i, (of, runs, seq) in enumerate(frags):
        assert sum(n for _, n in runs) == len(seq) - k + 1, f"frag {i}: runs do not cover the k-mers"
        if of & OPEN_LEFT:
            ends.setdefault(canon(seq[:k]), []).append((i, 0))
        if of & OPEN_RIGHT:
            ends.setdefault(canon(seq[-k:]), []).append((i, 1))
    partner = {}
    for key, lst in ends.items():
        assert len(lst) <= 2, f"{len(lst)} open ends share a boundary k-mer"
        if len(lst) == 2:
            a, b = lst
            assert a[0] != b[0] or a[1] != b[1]
            partner[a], partner[b] = b, a

    used = [False] * len(frags)
    tigs = []

    def walk(start, enter_side):
        # orient `start` so that enter_side is its left end
        seq, cids, _ = _oriented(frags[start], enter_side == 1)
        used[start] = True
        cur, out_side = start, 1 - enter_side
        circular = False
        while (cur, out_side) in partner:
            nxt, nside = partner[(cur, out_side)]
            if used[nxt]:
                assert nxt == start and nside == enter_side, "chain re-enters a used fragment"
                # the last k-mer is the start's first k-mer: keep it once
                assert seq[-k:] == seq[:k]
                x = cids[0] if cids[0] != FOREIGN else cids[-1]
                cids = [x] + cids[1:-1]
                seq = seq[:-1]
                circular = True
                break
            s2, c2, _ = _oriented(frags[nxt], nside == 1)
            assert s2[:k] == seq[-k:], "joined ends do not share the boundary k-mer"
            x = cids[-1] if cids[-1] != FOREIGN else c2[0]
            assert cids[-1] == FOREIGN or c2[0] == FOREIGN or cids[-1] == c2[0]
            cids = cids[:-1] + [x] + c2[1:]
            seq = seq + s2[k:]
            used[nxt] = True
            cur, out_side = nxt, 1 - nside
        tigs.append((seq, cids, circular))

    for i in range(len(frags)):
        if used[i]:
            continue
        if (i, 0) not in partner:
            walk(i, 0)
        elif (i, 1) not in partner:
            walk(i, 1)
    for i in range(len(frags)):  # what is left are cycles of joined fragments
        if not used[i]:
            walk(i, 0)
    return tigs


def split_monochromatic(tigs, k):
    """Unitigs as (sequence, cid): cuts at cid changes; extremal foreign
    placeholders are dropped, interior ones are an error."""
    out = []
    for seq, cids, circular in tigs:
        lo, hi = 0, len(cids)
        if circular:
            assert FOREIGN not in cids
            n = len(cids)
            if len(set(cids)) == 1:
                out.append((seq, cids[0], True))
                continue
            # rotate the cycle to start at a colour change; cycle base j is seq[j % n]
            r = next(i for i in range(n) if cids[i] != cids[i - 1])
            cyc = seq[:n] * (2 + (k - 1) // n)
            seq, cids = cyc[r:r + n + k - 1], cids[r:] + cids[:r]
            hi = n
        if lo < hi and cids[lo] == FOREIGN:
            lo += 1
        if lo < hi and cids[hi - 1] == FOREIGN:
            hi -= 1
        assert FOREIGN not in cids[lo:hi], "interior foreign k-mer"
        i = lo
        while i < hi:
            j = i
            while j < hi and cids[j] == cids[i]:
                j += 1
            out.append((seq[i:j + k - 1], cids[i], False))
            i = j
    return out


def kmers_of(seq, k):
    return [seq[i:i + k] for i in range(len(seq) - k + 1)]


def color_map(colour_seqs, k):
    """canonical k-mer -> colour bitmask (verify.py:103-129)."""
    S = {}
    for c, seqs in enumerate(colour_seqs):
        for s in seqs:
            run = []
            for ch in s.upper() + b"N":
                if ch in b"ACGT":
                    run.append(ch)
                    continue
                r = bytes(run)
                for x in kmers_of(r, k):
                    key = canon(x)
                    S[key] = S.get(key, 0) | (1 << c)
                run = []
    return S


def truth_unitigs(S, k):
    """Unitigs by definition (verify.py:133-220) as a list of (seq, mask, circular)."""
    def succ(x):
        return [x[1:] + b for b in (b"A", b"C", b"G", b"T") if canon(x[1:] + b) in S]

    def pred(x):
        return [b + x[:-1] for b in (b"A", b"C", b"G", b"T") if canon(b + x[:-1]) in S]

    def fwd(x):
        nb = {canon(y): y for y in succ(x)}
        if len(nb) != 1:
            return None
        y = next(iter(nb.values()))
        if len({canon(z) for z in pred(y)}) != 1 or S[canon(x)] != S[canon(y)]:
            return None
        return y

    def back(x):
        nb = {canon(y): y for y in pred(x)}
        if len(nb) != 1:
            return None
        z = next(iter(nb.values()))
        if len({canon(w) for w in succ(z)}) != 1 or S[canon(x)] != S[canon(z)]:
            return None
        return z

    seen, out = set(), []
    for c0 in S:
        if c0 in seen:
            continue
        x, local = c0, {c0}
        circular = False
        while True:
            z = back(x)
            if z is None:
                break
            if canon(z) in local:
                circular = True
                break
            x = z
            local.add(canon(z))
        seq = x
        seen.add(canon(x))
        while True:
            y = fwd(x)
            if y is None:
                break
            if canon(y) in seen:
                circular = True
                break
            seq += y[-1:]
            seen.add(canon(y))
            x = y
        out.append((seq, S[canon(seq[:k])], circular))
    return out


def unitig_key(seq, k):
    """Orientation- and rotation-free identity of a unitig: its canonical k-mers."""
    return frozenset(canon(x) for x in kmers_of(seq, k))


def check_path(seq, k, S):
    """Every k-mer of seq is in S and consecutive k-mers overlap by k-1 (by construction)."""
    return all(canon(x) in S for x in kmers_of(seq, k))

"""The cdbg-build host driver (command line of src/main.cpp:13-74, configuration
rules of builder.hpp:550-649).

CPU: argument and configuration errors exit with status 1 before any device
work.  GPU: a run over FASTA files leaves the same phase-2 artifacts as the
oracle's phase 2 over the bucket files the run kept (--keep-tmp), and the
resident mode (no bucket files) and the host stitch (--host-stitch) leave
identical deliverables.
"""
import os
import subprocess

import pytest

import oracle as O
from conftest import PKG_DIR, cdbg_gpu as G

CLI = os.path.join(PKG_DIR, "bin", "cdbg-build")
needs_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="cdbg-build not built")


def _run(*args):
    return subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=600)


def _list(tmp_path, colors=3, glen=15000, seed=9):
    data, off = G.synth(seed=seed, num_colors=colors, genome_len=glen, snp_rate=0.01, indel_rate=0.001)
    paths = []
    for c in range(colors):
        p = tmp_path / f"g{c}.fa"
        seq = bytes(data[off[c]:off[c + 1] - 1])
        p.write_bytes(b">g\n" + b"\n".join(seq[i:i + 80] for i in range(0, len(seq), 80)) + b"\n")
        paths.append(str(p))
    lst = tmp_path / "list.txt"
    lst.write_text("\n".join(paths) + "\n")
    return str(lst), paths


@needs_cli
@pytest.mark.parametrize("args,msg", [
    ((), "usage"),
    (("-i", "x", "-o", "y"), "usage"),
    (("-i", "x", "-o", "y", "-k", "64"), "k must satisfy"),
    (("-i", "x", "-o", "y", "-k", "31", "-m", "31"), "invalid m"),
    (("-i", "x", "-o", "y", "-k", "31", "-b", "17"), "-b must be in"),
    (("-i", "x", "-o", "y", "-k", "31", "-g", "2"), "-g must be at least"),
    (("-i", "x", "-o", "y", "-k", "31", "-g", "4", "-t", "10000"), "Reduce -t"),
    (("-i", "/nonexistent/list", "-o", "y", "-k", "31"), "cannot open filenames list"),
    (("-i", "x", "-o", "y", "-k", "31", "--bogus"), "unknown option"),
])
def test_cli_rejects(args, msg):
    r = _run(*args)
    assert r.returncode == 1
    assert msg in r.stderr


@needs_cli
def test_cli_rejects_nonempty_tmp_dir(tmp_path):
    lst, _ = _list(tmp_path, colors=1, glen=200)
    d = tmp_path / "scratch"
    d.mkdir()
    (d / "junk").write_text("x")
    r = _run("-i", lst, "-o", tmp_path / "out", "-k", 31, "-d", d)
    assert r.returncode == 1 and "is not empty" in r.stderr


@needs_cli
def test_cli_clamps_auto_bucket_count(tmp_path):
    # -g 100 at -t 1 sizes ~1.2e5 buckets by the reference's model
    # (builder.hpp:608-621); phase 1 keeps 16-bit bucket ids, so the count is
    # clamped to 65536 with a warning instead of failing after the input is read
    lst, _ = _list(tmp_path, colors=1, glen=200)
    r = _run("-i", lst, "-o", tmp_path / "out", "-k", 31, "-g", 100, "-t", 1, "-d", tmp_path / "s")
    assert "clamped to 65536" in r.stderr
    assert "num_buckets = 65536" in r.stdout


@pytest.mark.gpu
@needs_cli
@pytest.mark.parametrize("k,m,b", [(31, 19, 6), (63, 27, 5)])
def test_cli_phase2_only(tmp_path, k, m, b):
    """--phase2-only leaves the artifacts the reference's own stitch and emit
    read: frag_unitigs.bin, the placeholder + complete words of
    .color_sets, its .tmp_offsets sidecar, and the dictionary state."""
    lst, paths = _list(tmp_path)
    d = tmp_path / "scratch"
    out = tmp_path / "out"
    r = _run("-i", lst, "-o", out, "-k", k, "-m", m, "-b", b, "-d", d, "-t", 4, "--keep-tmp", "--phase2-only")
    assert r.returncode == 0, r.stderr
    B = 1 << b
    ob = O.Build.from_bucket_files(str(d), B, k, len(paths))
    assert (d / "frag_unitigs.bin").read_bytes() == ob.frags().tobytes()
    words, offs = ob.color_sets()
    full = ob.stat("cs_bits") // 64
    assert (tmp_path / "out.color_sets").read_bytes() == bytes(28) + words[:full].tobytes()
    assert (tmp_path / "out.color_sets.tmp_offsets").read_bytes() == offs.tobytes()
    meta = dict(line.split("=", 1) for line in (tmp_path / "out.phase2.txt").read_text().split())
    assert int(meta["num_kmers"]) == ob.stat("num_kmers")
    assert int(meta["num_color_sets"]) == ob.stat("num_sets")

    d2, out2 = tmp_path / "scratch2", tmp_path / "res"
    r = _run("-i", lst, "-o", out2, "-k", k, "-m", m, "-b", b, "-d", d2, "--resident", "--phase2-only")
    assert r.returncode == 0, r.stderr
    assert (d2 / "frag_unitigs.bin").read_bytes() == (d / "frag_unitigs.bin").read_bytes()
    assert (tmp_path / "res.color_sets").read_bytes() == (tmp_path / "out.color_sets").read_bytes()
    assert (tmp_path / "res.color_sets.tmp_offsets").read_bytes() == offs.tobytes()


def check_final_outputs(out, d, ob, k, num_colors):
    """The four deliverables of a full run against the oracle's phase 2 over
    the bucket files the run kept: .fa = the host stitch of the oracle's
    fragments, ascending cid; .u2c one bit per unitig, one set bit per colour
    set; .color_sets header, every stream word and the Elias-Fano offsets;
    metadata fields."""
    import struct
    import numpy as np
    from test_stitch_emit import read_elias_fano
    with G.stitch(ob.frags(), k) as U:
        want = U.unitigs()
    lines = open(f"{out}.fa", "rb").read().split(b"\n")
    fa = [(lines[2 * i + 1], int(lines[2 * i][1:])) for i in range(len(lines) // 2)]
    assert fa == want
    raw = open(f"{out}.u2c", "rb").read()
    nbits, nwords = struct.unpack_from("<QQ", raw)
    assert nbits == len(want)
    u2c = np.frombuffer(raw[16:], np.uint64)
    assert sum(bin(int(w)).count("1") for w in u2c) == ob.stat("num_sets")
    buf = open(f"{out}.color_sets", "rb").read()
    words, offs = ob.color_sets()
    nw = (ob.stat("cs_bits") + 63) // 64
    assert struct.unpack_from("<IIIQQ", buf) == (num_colors, num_colors // 4, 3 * num_colors // 4,
                                                 ob.stat("cs_bits"), nw)
    assert buf[28:28 + 8 * nw] == words[:nw].tobytes()
    vals, end = read_elias_fano(buf, 28 + 8 * nw)
    assert end == len(buf) and vals == [int(x) for x in offs] + [ob.stat("cs_bits")]
    assert not os.path.exists(f"{out}.color_sets.tmp_offsets")
    meta = dict(x.split("=") for x in open(f"{out}.metadata.txt").read().split())
    assert meta == {"k": str(k), "num_kmers": str(ob.stat("num_kmers")), "num_colors": str(num_colors),
                    "num_unitigs": str(len(want)), "num_color_sets": str(ob.stat("num_sets"))}
    return len(want)


@pytest.mark.gpu
@needs_cli
@pytest.mark.parametrize("k,m,b", [(31, 19, 6), (63, 27, 5)])
def test_cli_end_to_end(tmp_path, k, m, b):
    lst, paths = _list(tmp_path)
    d = tmp_path / "scratch"
    out = tmp_path / "out"
    r = _run("-i", lst, "-o", out, "-k", k, "-m", m, "-b", b, "-d", d, "-t", 4, "--keep-tmp")
    assert r.returncode == 0, r.stderr
    ob = O.Build.from_bucket_files(str(d), 1 << b, k, len(paths))
    assert (d / "frag_unitigs.bin").read_bytes() == ob.frags().tobytes()
    n = check_final_outputs(out, d, ob, k, len(paths))
    assert f"num_unitigs = {n}" in r.stdout
    # resident mode: the same deliverables
    d2, out2 = tmp_path / "scratch2", tmp_path / "res"
    r = _run("-i", lst, "-o", out2, "-k", k, "-m", m, "-b", b, "-d", d2, "--resident")
    assert r.returncode == 0, r.stderr
    for ext in (".fa", ".u2c", ".color_sets", ".metadata.txt"):
        assert open(f"{out2}{ext}", "rb").read() == open(f"{out}{ext}", "rb").read(), ext
    # the host stitch (--host-stitch) writes the same deliverables as the device stitch
    d3, out3 = tmp_path / "scratch3", tmp_path / "hst"
    r = _run("-i", lst, "-o", out3, "-k", k, "-m", m, "-b", b, "-d", d3, "--host-stitch")
    assert r.returncode == 0, r.stderr
    for ext in (".fa", ".u2c", ".color_sets", ".metadata.txt"):
        assert open(f"{out3}{ext}", "rb").read() == open(f"{out}{ext}", "rb").read(), ext

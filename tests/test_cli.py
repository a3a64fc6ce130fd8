"""The command line (sdhash_amd/sdhash-amd) and a caller written against the
reference's class interface (tests/caller_compat.cpp, built against
include/ under the reference's header names).

CPU: the caller compiles and links against libsdbf_amd.so (the interface is
source-compatible with the reference's call sites).  GPU: both programs run,
and their printed digests and comparison lines equal what the oracle and the
reference's formatting give (sdbf_class.cc:294-334, sdbf_set.cc:240-394,
segment naming and sizes of sdhash_threads.cc:300-373)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from sdhash_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sdhash_amd")
CLI = os.path.join(PKG, "sdhash-amd")
KIB = 1024
MIB = 1 << 20


def build_caller(out):
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "caller_compat.cpp"), "-o", out, "-L", PKG, "-l:libsdbf_amd.so",
                    "-l:libsdbfx.so", "-Wl,-rpath," + PKG], check=True)
    return out


def test_caller_compiles_against_the_interface(tmp_path):
    if not os.path.exists(os.path.join(PKG, "libsdbf_amd.so")):
        from sdhash_amd import build
        build.build()
    assert os.path.exists(build_caller(str(tmp_path / "caller_compat")))


def run(args, cwd, stdin=None):
    r = subprocess.run(args, cwd=cwd, capture_output=True, timeout=300, input=stdin)
    assert r.returncode == 0, (args, r.returncode, r.stderr[-2000:])
    return r.stdout


def lines(text):
    return sorted(l for l in text.decode().splitlines() if l)


def compare_lines(pairs, sep="|"):
    """Printed lines of sdbf_set::compare_* (sdbf_set.cc:256-262): setw(3)
    zero-filled scores, -1 unpadded; scores below the threshold 1 omitted."""
    out = []
    for a, b, sc in pairs:
        if sc >= 1:
            out.append("%s%s%s%s%s" % (a, sep, b, sep, "%03d" % sc if sc != -1 else "-1"))
    return sorted(out)


@pytest.mark.gpu
def test_cli_segments_tail_and_sizes(tmp_path):
    # 2 MiB + 100 bytes in 1 MiB segments at -b 16: the 100-byte tail joins the
    # second segment, whose digest covers 1 MiB + 200 bytes (the tail twice,
    # zero-filled) and whose recorded size is the 1 MiB + 100 bytes read
    data = synth.disk_image(301, 2 * MIB + 100)
    (tmp_path / "img.bin").write_bytes(data.tobytes())
    out = run([CLI, "-b", "16", "-z", "1", "img.bin"], tmp_path).decode()
    seg0 = oracle.digest_block(data[:MIB], 16384, b"img.bin.0000M")
    seg0.orig_size = MIB
    tail = np.concatenate([data[MIB:], np.zeros(100, np.uint8)])
    seg1 = oracle.digest_block(tail, 16384, b"img.bin.0001M")
    seg1.orig_size = MIB + 100
    assert out == seg0.to_string() + seg1.to_string()
    # the same through the class interface's constructors
    caller = build_caller(str(tmp_path / "caller_compat"))
    assert run([caller, "hash", "16", "1", "img.bin"], tmp_path).decode() == out


@pytest.mark.gpu
def test_cli_hash_and_compare(tmp_path):
    files = synth.file_set(302, 6, 300 * KIB)
    names = ["f%d.bin" % i for i in range(len(files))]
    for n, f in zip(names, files):
        (tmp_path / n).write_bytes(f.tobytes())
    ods = [oracle.digest_stream(f, n.encode()) for f, n in zip(files, names)]
    want_text = "".join(d.to_string() for d in ods)
    # stream digests, printed
    assert run([CLI] + names, tmp_path).decode() == want_text
    caller = build_caller(str(tmp_path / "caller_compat"))
    assert run([caller, "hash", "0", "128"] + names, tmp_path).decode() == want_text
    # -g: all pairs
    sc = oracle.compare_all(ods)
    pairs, k = [], 0
    for i in range(len(ods)):
        for j in range(i + 1, len(ods)):
            pairs.append((names[i], names[j], int(sc[k])))
            k += 1
    assert (sc >= 1).sum() > 3
    assert lines(run([CLI, "-g"] + names, tmp_path)) == compare_lines(pairs)
    # -o then -c on the saved set, and the caller's compare_all_quiet
    run([CLI, "-o", "set"] + names, tmp_path)
    assert (tmp_path / "set.sdbf").read_text() == want_text
    assert lines(run([CLI, "-c", "set.sdbf"], tmp_path)) == compare_lines(pairs)
    assert lines(run([caller, "all", "1", "0", "set.sdbf"], tmp_path)) == compare_lines(pairs)
    # -c q t with sampling (-s), and the caller's compare_to_quiet
    q = [oracle.digest_stream(f[o:o + 60 * KIB], b"q%d" % i)
         for i, (f, o) in enumerate(zip(files, (0, 9000, 77000, 130000, 200000, 5)))]
    (tmp_path / "q.sdbf").write_text("".join(d.to_string() for d in q))
    for smp in (0, 3):
        sq = oracle.compare_to(q, ods, sample=smp)
        want = compare_lines([(q[i].name.decode(), names[j], int(sq[i, j]))
                              for i in range(len(q)) for j in range(len(ods))])
        assert lines(run([CLI, "-c", "q.sdbf", "set.sdbf", "-s", str(smp)], tmp_path)) == want, smp
        assert lines(run([caller, "two", "1", str(smp), "0", "q.sdbf", "set.sdbf"], tmp_path)) == want, smp
    # --fast on -c q t: the query set folded only (sdbf_set.cc:367-376)
    sq = oracle.compare_to([d.folded() for d in q], ods)
    want = compare_lines([(q[i].name.decode(), names[j], int(sq[i, j]))
                          for i in range(len(q)) for j in range(len(ods))])
    assert lines(run([CLI, "-c", "q.sdbf", "set.sdbf", "--fast"], tmp_path)) == want
    assert lines(run([caller, "two", "1", "0", "1", "q.sdbf", "set.sdbf"], tmp_path)) == want
    # stdin with -z 0 (segments of 128 MiB, sdhash.cc:461-463) and the class API checks
    out = run([CLI, "-z", "0", "-"], tmp_path, stdin=files[0].tobytes()).decode()
    d = oracle.digest_block(files[0], 16384, b"stdin.0000M")
    assert out == d.to_string()
    assert run([caller, "api"], tmp_path).decode().startswith("api ok")


def idx_bits(caller, path, tmp_path):
    out = subprocess.run([caller, "idxr", path, str(tmp_path / "idx.bin")], capture_output=True, text=True,
                         check=True).stdout.split()
    return np.fromfile(str(tmp_path / "idx.bin"), np.uint8), int(out[1]), out[3]


@pytest.mark.gpu
def test_cli_index_and_index_search(tmp_path):
    # --index (sdhash_threads.cc:376-492): small files stream-hashed in order
    # against one 64 MiB index -> out0.sdbf and out0.sdbf.idx; a large file
    # in 16 KiB blocks -> big0.sdbf(.idx); then --index-search of block-mode
    # queries against the directory (sdhash.cc:241-264, 640-652)
    caller = build_caller(str(tmp_path / "caller_compat"))
    img = synth.disk_image(310, 3 * MIB)
    rng = np.random.Generator(np.random.PCG64(311))
    small = [img[:700 * KIB], img[500 * KIB:1300 * KIB], np.concatenate([img[:200 * KIB], synth.segment_mix(rng, 300 * KIB)]),
             img[2 * MIB:]]
    names = ["s%d.bin" % k for k in range(len(small))]
    for n, f in zip(names, small):
        (tmp_path / n).write_bytes(f.tobytes())
    big = synth.disk_image(312, 17 * MIB)
    (tmp_path / "big.bin").write_bytes(big.tobytes())
    (tmp_path / "ix").mkdir()
    run([CLI, "--index", "-o", "ix/out"] + names, tmp_path)
    oix = oracle.Index(64 << 20)
    want = "".join(oracle.digest_stream_indexed(f, oix, n.encode()).to_string() for f, n in zip(small, names))
    assert (tmp_path / "ix" / "out0.sdbf").read_text() == want
    bits, ec, nm = idx_bits(caller, str(tmp_path / "ix" / "out0.sdbf.idx"), tmp_path)
    assert ec == oix.elem_count and nm == "ix/out0.sdbf" and np.array_equal(bits, oix.bf)
    run([CLI, "--index", "-o", "big"] + ["big.bin"], tmp_path)
    bix = oracle.Index(64 << 20)
    d = oracle.digest_block_indexed(big, 16384, b"big.bin", index=bix)[0]
    assert (tmp_path / "big0.sdbf").read_text() == d.to_string()
    bits, ec, _ = idx_bits(caller, str(tmp_path / "big0.sdbf.idx"), tmp_path)
    assert ec == bix.elem_count and np.array_equal(bits, bix.bf)
    # search: queries cut from the indexed data, and fresh data.  (s1 and
    # s2 begin with data that is already indexed: the gate skips those
    # features after they set their filter bits, so the filter fills up and
    # nothing after it reaches the index, as in the reference.)
    qs = [img[100 * KIB:600 * KIB], img[2 * MIB + 100 * KIB:2 * MIB + 600 * KIB], synth.segment_mix(rng, 200 * KIB)]
    qn = ["q%d.bin" % k for k in range(len(qs))]
    for n, q in zip(qn, qs):
        (tmp_path / n).write_bytes(q.tobytes())
    r = subprocess.run([CLI, "--index-search", "ix"] + qn, cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = []
    for n, q in zip(qn, qs):
        _, m = oracle.digest_block_indexed(q, 16384, n.encode(), search=[oix])
        lines += ["%s [%d] |ix/out0.sdbf|%d" % (n, k, c) for k, c in enumerate(m[:, 0]) if c > 4]
    assert len(lines) > 10
    assert r.stderr.decode().splitlines() == lines
    run([CLI, "--index-search", "ix", "-o", "res"] + qn, tmp_path)
    assert (tmp_path / "res.idx-result").read_text().splitlines() == lines

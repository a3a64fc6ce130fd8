"""The sdhash-server request handler (include/sdhash_srv.h,
sdhash-server/src/sdhash-srv-main.cc:52-660) on the device, driven by
tests/srv_driver.cpp: set loading with indexes, compareTwo / compareAll
results, hashString into a saved set and its index search, result status,
against the oracle and the reference's output formats."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from sdhash_amd import synth

from test_cli import CLI, ROOT, PKG, compare_lines, run

KIB = 1024
MIB = 1 << 20


def build_driver(out):
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "srv_driver.cpp"), "-o", out, "-L", PKG, "-l:libsdbf_amd.so",
                    "-l:libsdbfx.so", "-Wl,-rpath," + PKG], check=True)
    return out


def test_driver_compiles(tmp_path):
    if not os.path.exists(os.path.join(PKG, "libsdbf_amd.so")):
        from sdhash_amd import build
        build.build()
    assert os.path.exists(build_driver(str(tmp_path / "srv_driver")))


def sections(text):
    out, cur, buf = {}, None, []
    for line in text.splitlines():
        if line.startswith("== ") and line != "== end":
            cur, buf = line[3:], []
        elif line == "== end":
            out[cur] = buf
            cur = None
        elif cur is not None:
            buf.append(line)
    return out


@pytest.mark.gpu
def test_handler_requests(tmp_path):
    driver = build_driver(str(tmp_path / "srv_driver"))
    home, src = tmp_path / "home", tmp_path / "src"
    home.mkdir()
    src.mkdir()
    img = synth.disk_image(501, 4 * MIB)
    rng = np.random.Generator(np.random.PCG64(502))
    # set "a": stream digests indexed by sdhash --index; set "b": plain
    af = [img[:600 * KIB], img[1 * MIB:1 * MIB + 500 * KIB], img[2 * MIB:3 * MIB]]
    an = ["src/a%d.bin" % k for k in range(len(af))]
    bf = [img[100 * KIB:400 * KIB], synth.segment_mix(rng, 300 * KIB), img[2 * MIB + 50 * KIB:2 * MIB + 900 * KIB]]
    bn = ["src/b%d.bin" % k for k in range(len(bf))]
    for n, f in zip(an + bn, af + bf):
        (tmp_path / n).write_bytes(f.tobytes())
    run([CLI, "--index", "-o", "home/a"] + an, tmp_path)
    os.rename(home / "a0.sdbf", home / "a.sdbf")
    os.rename(home / "a0.sdbf.idx", home / "a.sdbf.idx")
    oix = oracle.Index(64 << 20)
    ao = [oracle.digest_stream_indexed(f, oix, n.encode()) for f, n in zip(af, an)]
    bo = [oracle.digest_stream(f, n.encode()) for f, n in zip(bf, bn)]
    (home / "b.sdbf").write_text("".join(d.to_string() for d in bo))
    qf = [img[150 * KIB:550 * KIB], img[2 * MIB + 100 * KIB:2 * MIB + 400 * KIB]]
    qn = ["q%d.bin" % k for k in range(len(qf))]
    for n, f in zip(qn, qf):
        (tmp_path / n).write_bytes(f.tobytes())
    r = subprocess.run([driver, "home", "src"] + qn, cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    s = sections(r.stdout.decode())
    # sets 1 = a, 2 = b (home listed in name order); the .idx file itself
    # does not parse as a set
    assert [l.split()[:2] for l in s["setList"]] == [["1", "a"], ["2", "b"]]
    assert s["setList json"][0].startswith('{ "aaData": [["1","a","')
    assert [l.split()[0] for l in s["displaySet 2"]] == bn
    sc = oracle.compare_to(ao, bo)
    want = compare_lines([(ao[i].name.decode(), bo[j].name.decode(), int(sc[i, j]))
                          for i in range(len(ao)) for j in range(len(bo))])
    assert sorted(s["compareTwo"]) == want and len(want) > 2
    assert s["status"] == ["complete (%d)" % len(want)]
    assert s["info"] == ["a -- b"]
    sc = oracle.compare_all(bo)
    pairs, k = [], 0
    for i in range(len(bo)):
        for j in range(i + 1, len(bo)):
            pairs.append((bo[i].name.decode(), bo[j].name.decode(), int(sc[k])))
            k += 1
    assert sorted(s["compareAll"]) == compare_lines(pairs)
    assert len(s["resultsList"]) == 2
    # hashString(0): stream digests of the query files, saved as fresh.sdbf
    qo = [oracle.digest_stream(f, n.encode()) for f, n in zip(qf, qn)]
    text = "".join(d.to_string() for d in qo)
    assert "\n".join(s["hashString"]) + "\n" == text
    assert (home / "fresh.sdbf").read_text() == text
    assert s["name"] == ["fresh"]
    # hashString(3): 16 KiB blocks looked up in the index of set a
    lines = []
    for n, f in zip(qn, qf):
        _, m = oracle.digest_block_indexed(f, 16384, n.encode(), search=[oix])
        lines += ["%s [%d] |a|%d" % (n, k, c) for k, c in enumerate(m[:, 0]) if c > 4]
    assert len(lines) > 5 and s["search"] == lines
    assert sorted(s["sources"]) == sorted(an + bn)
    assert s["removed"] == ["not found"]

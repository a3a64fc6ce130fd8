"""The C++ host layer's text form: parsing the oracle's sdbf text with
sdbf_set(const char*) and writing it back with sdbf::to_string is
byte-identical (sdbf/sdbf_class.cc:144-215, 294-334).  CPU only."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from sdhash_amd import build, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def roundtrip(tmp_path_factory):
    if not os.path.exists(build.CXXLIB):
        build.build()
    exe = str(tmp_path_factory.mktemp("codec") / "codec_roundtrip")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "codec_roundtrip.cpp"), "-o", exe,
                    "-L", os.path.dirname(build.CXXLIB), "-l:libsdbf_amd.so", "-l:libsdbfx.so",
                    "-Wl,-rpath," + os.path.dirname(build.CXXLIB)], check=True)
    return exe


def test_roundtrip_stream_and_block(roundtrip, tmp_path):
    files = synth.file_set(41, 5, 150 * 1024)
    ds = [oracle.digest_stream(f, b"dir/file%d.bin" % i) for i, f in enumerate(files[:3])]
    ds += [oracle.digest_block(f, 16384, b"img%d" % i) for i, f in enumerate(files[3:])]
    ds.append(oracle.digest_stream(np.zeros(600, np.uint8), b"zeros"))
    text = "".join(d.to_string() for d in ds)
    p = tmp_path / "set.sdbf"
    p.write_text(text)
    r = subprocess.run([roundtrip, str(p)], capture_output=True, text=True, check=True)
    assert r.stdout == text
    n, nf, size = map(int, r.stderr.split())
    assert n == len(ds) and nf == sum(d.bf_count for d in ds)
    assert size == sum(d.orig_size for d in ds)


def test_invalid_text_is_format_error(roundtrip, tmp_path):
    d = oracle.digest_block(synth.file_set(42, 1, 40000)[0], 16384, b"x")
    bad = d.to_string().replace(":", ":!", 20)   # corrupt a base64 field
    p = tmp_path / "bad.sdbf"
    p.write_text(bad)
    r = subprocess.run([roundtrip, str(p)], capture_output=True, text=True)
    assert r.returncode == 3 and "error -2" in r.stderr

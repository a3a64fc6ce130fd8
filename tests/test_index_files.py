"""Index files (.idx, bloom_filter::write_out / bloom_filter(string),
sdbf/bloom_filter.cc:77-114, 176-228) through the class interface
(tests/caller_compat.cpp against include/): the header fields, and the bits
after an LZ4 round trip, against the oracle's index after the same inserts.
Where the reference sdhash tree is present, its own LZ4 (lz4/lz4.cc, built
into oracle/_ref/lz4check by oracle/Makefile) must decode the files written
here, and the reader here must decode what the reference's compressor
writes.  CPU only (no device calls)."""
import os
import subprocess

import numpy as np
import pytest

import oracle

from test_cli import build_caller

M64 = (1 << 64) - 1


def xorshift_words(seed, n):
    x = seed
    out = np.zeros((n, 5), np.uint32)
    for k in range(n):
        for j in range(5):
            x ^= (x << 13) & M64
            x ^= x >> 7
            x ^= (x << 17) & M64
            out[k, j] = x & 0xFFFFFFFF
    return out


def header(path):
    raw = open(path, "rb").read()
    nl = raw.index(b"\n")
    f = raw[:nl].decode().split(":", 6)
    return f, raw[nl + 1:]


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    return build_caller(str(tmp_path_factory.mktemp("idx") / "caller_compat"))


@pytest.mark.parametrize("size,n", [(4096, 300), (65536, 3000), (64 << 20, 2000)])
def test_index_file_round_trip(caller, tmp_path, size, n):
    path = str(tmp_path / "a.idx")
    r = subprocess.run([caller, "idxw", str(size), str(n), "7", path], capture_output=True, text=True, check=True)
    ref = oracle.Index(size)
    for h in xorshift_words(7, n):
        ref.insert(h)
    assert int(r.stdout) == ref.elem_count
    f, payload = header(path)
    assert f[:6] == ["sdbf-idx", str(size), str(ref.elem_count), "5", str(8 * size - 1), str(len(payload))]
    assert f[6] == "set.sdbf"
    if size == 64 << 20:
        assert len(payload) < size // 8   # mostly empty: compressed
    out = subprocess.run([caller, "idxr", path, str(tmp_path / "b.bin")], capture_output=True, text=True,
                         check=True).stdout.split()
    assert out == [str(size), str(ref.elem_count), str(8 * size - 1), "set.sdbf"]
    assert np.array_equal(np.fromfile(str(tmp_path / "b.bin"), np.uint8), ref.bf)
    if os.path.exists(oracle.REF_LZ4):
        (tmp_path / "p.lz4").write_bytes(payload)
        subprocess.run([oracle.REF_LZ4, "d", str(tmp_path / "p.lz4"), str(tmp_path / "r.bin"), str(size)], check=True)
        assert np.array_equal(np.fromfile(str(tmp_path / "r.bin"), np.uint8), ref.bf)


def test_index_file_from_reference_compressor(caller, tmp_path):
    if not os.path.exists(oracle.REF_LZ4):
        pytest.skip("the reference sdhash tree is not present")
    rng = np.random.Generator(np.random.PCG64(9))
    bits = np.zeros(1 << 20, np.uint8)
    bits[rng.integers(0, bits.size, 30000)] = rng.integers(1, 256, 30000, dtype=np.uint8)
    bits[: 5000] = rng.integers(0, 256, 5000, dtype=np.uint8)
    bits.tofile(str(tmp_path / "bits.bin"))
    subprocess.run([oracle.REF_LZ4, "c", str(tmp_path / "bits.bin"), str(tmp_path / "c.lz4")], check=True)
    payload = (tmp_path / "c.lz4").read_bytes()
    hdr = "sdbf-idx:%d:%d:5:%d:%d:ref.sdbf\n" % (bits.size, 777, 8 * bits.size - 1, len(payload))
    (tmp_path / "r.idx").write_bytes(hdr.encode() + payload)
    out = subprocess.run([caller, "idxr", str(tmp_path / "r.idx"), str(tmp_path / "o.bin")], capture_output=True,
                         text=True, check=True).stdout.split()
    assert out == [str(bits.size), "777", str(8 * bits.size - 1), "ref.sdbf"]
    assert np.array_equal(np.fromfile(str(tmp_path / "o.bin"), np.uint8), bits)
    # a truncated payload does not decode: -2 (bloom_filter.cc:106-108)
    (tmp_path / "t.idx").write_bytes(hdr.encode() + payload[:-40])
    out = subprocess.run([caller, "idxr", str(tmp_path / "t.idx"), str(tmp_path / "o2.bin")], capture_output=True,
                         text=True, check=True).stdout
    assert out.startswith("throw -2")

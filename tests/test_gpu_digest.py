"""Digest parity on the GPU: filters, element counts and Hamming weights of
every digest are bit-identical with the CPU oracle."""
import numpy as np
import pytest

import oracle
import sdhash_amd
from sdhash_amd import synth

pytestmark = pytest.mark.gpu

KIB = 1024
MIB = 1 << 20


def check_same(gset, i, od, label):
    f, ec, hm = gset.read(i)
    info = gset.info(i)
    assert info.bf_count == od.bf_count, (label, info.bf_count, od.bf_count)
    if od.elem_counts is None:
        assert info.last_count == od.last_count, label
    bad = np.flatnonzero(f != od.filters)
    assert bad.size == 0, (label, "first differing filter", bad[0] // 256 if bad.size else None)
    od.cstruct()
    assert np.array_equal(hm, od.hamming[: od.bf_count]), label
    if od.elem_counts is not None:
        assert np.array_equal(ec, od.elem_counts), label
    else:
        want = np.full(od.bf_count, 160, np.uint16)
        want[-1] = od.last_count
        assert np.array_equal(ec, want), label


def test_block_mode_parity(engine):
    img = synth.disk_image(21, 4 * MIB + 16 * KIB * 3 + 700)
    pieces = [img, img[: 16 * KIB * 5 + 300], img[:512], img[:513], img[: 16 * KIB - 1],
              img[100: 100 + 40000]]
    print("block sizes", [p.size for p in pieces], flush=True)
    s = engine.new_set()
    st = s.digest(pieces, dd_block=16384)
    assert np.all(st == 0)
    for i, p in enumerate(pieces):
        check_same(s, i, oracle.digest_block(p, 16384), f"block {i} size {p.size}")


def test_block_mode_small_dd(engine):
    img = synth.disk_image(22, 300 * KIB + 1000)
    s = engine.new_set()
    assert np.all(s.digest([img], dd_block=4096) == 0)
    check_same(s, 0, oracle.digest_block(img, 4096), "dd 4096")


def test_stream_mode_parity_sizes(engine):
    rng = np.random.Generator(np.random.PCG64(23))
    sizes = [512, 600, 1000, 4 * KIB, 16 * KIB + 64, 16 * KIB + 65, 100 * KIB, MIB, 3 * MIB + 17]
    bufs = [synth.segment_mix(rng, n) for n in sizes]
    print("stream sizes", sizes, flush=True)
    s = engine.new_set()
    st = s.digest(bufs)
    assert np.all(st == 0)
    for i, b in enumerate(bufs):
        check_same(s, i, oracle.digest_stream(b), f"stream size {b.size}")


def test_stream_mode_planted_set(engine):
    files = synth.file_set(24, 16, MIB)
    s = engine.new_set()
    assert np.all(s.digest(files) == 0)
    for i, b in enumerate(files):
        check_same(s, i, oracle.digest_stream(b), f"file {i}")


def test_stream_mode_multi_chunk(engine):
    # 32 MiB restart boundary: two units, de-duplication count reset between
    b = synth.disk_image(25, 32 * MIB + 3 * MIB + 5)
    print("stream 35 MiB", flush=True)
    s = engine.new_set()
    assert np.all(s.digest([b]) == 0)
    check_same(s, 0, oracle.digest_stream(b), "35 MiB")


def test_too_small_inputs_are_skipped(engine):
    s = engine.new_set()
    st = s.digest([np.zeros(100, np.uint8), np.zeros(511, np.uint8), np.ones(512, np.uint8)])
    assert list(st) == [-3, -3, 0]
    assert len(s) == 1


def test_degenerate_inputs(engine):
    bufs = [np.zeros(100000, np.uint8), np.full(70000, 7, np.uint8),
            np.tile(np.arange(64, dtype=np.uint8), 2000), np.tile(np.arange(16, dtype=np.uint8), 5000)]
    s = engine.new_set()
    assert np.all(s.digest(bufs) == 0)
    for i, b in enumerate(bufs):
        check_same(s, i, oracle.digest_stream(b), f"degenerate {i}")
    t = engine.new_set()
    assert np.all(t.digest(bufs, dd_block=16384) == 0)
    for i, b in enumerate(bufs):
        check_same(t, i, oracle.digest_block(b, 16384), f"degenerate block {i}")


def test_device_resident_input(engine):
    files = synth.file_set(26, 4, 300 * KIB)
    offs = []
    pos = 0
    for f in files:
        offs.append(pos)
        pos += (f.size + 255) // 256 * 256
    buf = engine.alloc(pos)
    for o, f in zip(offs, files):
        buf.upload(o, f)
    s = engine.new_set()
    st = s.digest_device(buf, offs, [f.size for f in files], dd_block=[0, 16384, 0, 16384])
    assert np.all(st == 0)
    for i, f in enumerate(files):
        od = oracle.digest_block(f, 16384) if i % 2 else oracle.digest_stream(f)
        check_same(s, i, od, f"device {i}")
    buf.free()


def test_set_text_form(engine):
    # sdbf_set::to_string of a mixed stream / block-mode set, the base64
    # written by the device: byte-identical with the oracle's text
    files = synth.file_set(61, 6, 300 * KIB)
    files.append(synth.disk_image(62, 3 * MIB + 700))
    names = [b"dir/f%d.bin" % i for i in range(len(files))]
    modes = [0, 16384, 0, 4096, 0, 16384, 16384]
    s = engine.new_set()
    assert np.all(s.digest(files, names, dd_block=modes) == 0)
    ods = [oracle.digest_block(f, m, n) if m else oracle.digest_stream(f, n)
           for f, n, m in zip(files, names, modes)]
    want = "".join(od.to_string() for od in ods).encode()
    got = s.to_text()
    assert got == want, next(i for i in range(min(len(got), len(want))) if got[i] != want[i])
    assert engine.new_set().to_text() == b""


def test_set_add_text(engine):
    # the reverse: the oracle's text read into a set (sdbf_set's file reader),
    # the bodies decoded by the device; every filter, element count and
    # Hamming weight, the scores and the text written back
    files = synth.file_set(63, 40, 200 * KIB)
    files += [synth.disk_image(64, 20 * MIB + 123), synth.disk_image(65, 9 * MIB)]
    names = [b"in/%03d" % i for i in range(len(files))]
    modes = [16384 if i % 3 == 0 else 0 for i in range(len(files))]
    modes[-2:] = [0, 4096]
    ods = [oracle.digest_block(f, m, n) if m else oracle.digest_stream(f, n)
           for f, n, m in zip(files, names, modes)]
    text = "".join(od.to_string() for od in ods).encode()
    s = engine.new_set()
    assert s.add_text(text) == len(ods)
    for i, od in enumerate(ods):
        f, ec, hm = s.read(i)
        info = s.info(i)
        assert info.bf_count == od.bf_count and info.name.encode() == od.name, i
        assert info.orig_size == od.orig_size and info.dd_block_size == od.dd_block_size, i
        assert np.array_equal(f, od.filters), i
        od.cstruct()
        assert np.array_equal(hm, od.hamming[: od.bf_count]), i
        if od.elem_counts is not None:
            assert np.array_equal(ec, od.elem_counts), i
        else:
            assert info.last_count == od.last_count, i
    assert s.to_text() == text
    assert np.array_equal(engine.compare_all(s), oracle.compare_all(ods, threads=oracle.threads_available()))
    # appended after digests already in the set
    s2 = engine.new_set()
    assert np.all(s2.digest(files[:3], names[:3], dd_block=modes[:3]) == 0)
    assert s2.add_text(text) == len(ods)
    assert s2.to_text() == b"".join(od.to_string().encode() for od in ods[:3]) + text
    # a malformed body leaves the set as it was
    bad = text.replace(b"A", b"!", 1)
    with pytest.raises(sdhash_amd.SdbfxError):
        s2.add_text(bad)
    assert len(s2) == 3 + len(ods)

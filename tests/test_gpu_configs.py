"""Parity at the full sizes of the two reference-runnable workloads, against
the CPU oracle on every host thread:

- the plain all-pairs method: 64 files of 1 MiB of segment mix with 0-50 %
  planted cross-file chunks, stream digests, every one of the 2,016 scores
  (sdhash -g);
- block-mode search: a 1 GiB image hashed with -b 16 as eight 128 MiB
  segments, 4,096 fragments of 4 KiB (half cut from the image, half fresh),
  every digest byte and all 32,768 fragment x segment scores (sdhash -c).

Both comparison kernels are checked (the `cmp_engine` fixture)."""
import numpy as np
import pytest

import oracle
from sdhash_amd import synth

pytestmark = pytest.mark.gpu

KIB = 1024
MIB = 1 << 20
GIB = 1 << 30


def same_digest(gset, i, od):
    flt, ec, hm = gset.read(i)
    info = gset.info(i)
    assert info.bf_count == od.bf_count, (i, info.bf_count, od.bf_count)
    assert np.array_equal(flt, od.filters), i
    if od.elem_counts is not None:
        assert np.array_equal(ec, od.elem_counts), i
    else:
        assert info.last_count == od.last_count, i


@pytest.fixture(scope="module")
def plain_set():
    files = synth.file_set(101, 64, MIB)
    names = [b"dir/file%02d.bin" % i for i in range(64)]
    ods = oracle.digest_stream_many(files, names, threads=oracle.threads_available())
    return files, names, ods


def test_plain_all_pairs_64_files(cmp_engine, plain_set):
    files, names, ods = plain_set
    s = cmp_engine.new_set()
    assert np.all(s.digest(files, names) == 0)
    for i, od in enumerate(ods):
        same_digest(s, i, od)
    got = cmp_engine.compare_all(s)
    want = oracle.compare_all(ods, threads=oracle.threads_available())
    assert got.size == 64 * 63 // 2
    assert np.array_equal(got, want), np.flatnonzero(got != want)[:10]
    # planted chunks: the scores span the range
    assert want.max() >= 50 and len(set(want.tolist())) >= 10


@pytest.fixture(scope="module")
def search_set():
    th = oracle.threads_available()
    img = synth.disk_image(1, GIB)
    frags = synth.fragments(2, img, 4096, 4 * KIB)
    seg = 128 * MIB
    seg_names = [b"image.%04dM" % (i * 128) for i in range(8)]
    fr_names = [b"frag%04d" % i for i in range(4096)]
    t = [oracle.digest_block(img[i * seg:(i + 1) * seg], 16384, seg_names[i], threads=th) for i in range(8)]
    q = oracle.digest_stream_many(frags, fr_names, threads=th)
    return img, frags, seg_names, fr_names, t, q


def test_block_search_1gib(cmp_engine, search_set):
    img, frags, seg_names, fr_names, t, q = search_set
    seg = 128 * MIB
    dimg = cmp_engine.alloc(GIB)
    dimg.upload(0, img)
    gt = cmp_engine.new_set()
    assert np.all(gt.digest_device(dimg, [i * seg for i in range(8)], [seg] * 8, seg_names, dd_block=16384) == 0)
    dimg.free()
    for i, od in enumerate(t):
        same_digest(gt, i, od)
    gq = cmp_engine.new_set()
    assert np.all(gq.digest(frags, fr_names) == 0)
    for i, od in enumerate(q):
        same_digest(gq, i, od)
    got = cmp_engine.compare_to(gq, gt)
    want = oracle.compare_to(q, t, threads=oracle.threads_available())
    assert got.shape == (4096, 8)
    assert np.array_equal(got, want), np.argwhere(got != want)[:10]
    # the fragments cut from the image find their segment
    assert (want[:2048].max(axis=1) > 0).sum() > 1000

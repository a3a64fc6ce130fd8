"""Parity on the two batched workloads, against the CPU oracle on every host
thread:

- batched digesting: ragged files of log-normal size (median 256 KiB,
  sigma 1.5, clamped to 4 KiB-64 MiB) from the segment mix with 5 % planted
  cross-file chunks, digested in one call under the CLI's file policy
  (inputs of 16 MiB and more in 16 KiB block mode, sdhash.cc:554-618): every
  digest byte;
- batched all-pairs: -b 16 digests of 1-6,000 filters with planted shared
  chunks, all pairs, both comparison kernels: every score (and so every
  score >= 21, the sweep's threshold)."""
import numpy as np
import pytest

import oracle
from sdhash_amd import synth

pytestmark = pytest.mark.gpu

KIB = 1024
MIB = 1 << 20


def same_digest(gset, i, od):
    flt, ec, hm = gset.read(i)
    info = gset.info(i)
    assert info.bf_count == od.bf_count, (i, info.bf_count, od.bf_count)
    assert np.array_equal(flt, od.filters), i
    if od.elem_counts is not None:
        assert np.array_equal(ec, od.elem_counts), i
    else:
        assert info.last_count == od.last_count, i


def test_batched_lognormal_digests(engine):
    sizes = synth.lognormal_sizes(201, 300)
    sizes[:2] = [20 * MIB, 33 * MIB]   # the policy's block-mode branch, and a 32 MiB unit crossing
    files = synth.ragged_set(202, sizes, max_frac=0.05)
    names = [b"batch/f%04d" % i for i in range(len(files))]
    modes = [16384 if f.size >= 16 * MIB else 0 for f in files]
    s = engine.new_set()
    st = s.digest(files, names, dd_block=modes)
    assert np.all(st == 0)
    th = oracle.threads_available()
    stream_idx = [i for i, m in enumerate(modes) if m == 0]
    ods = dict(zip(stream_idx, oracle.digest_stream_many([files[i] for i in stream_idx],
                                                          [names[i] for i in stream_idx], threads=th)))
    for i, m in enumerate(modes):
        if m:
            ods[i] = oracle.digest_block(files[i], m, names[i], threads=th)
    for i in range(len(files)):
        same_digest(s, i, ods[i])
    assert sum(modes) > 0 and sum(files[i].size for i in range(len(files))) > 100 * MIB


def test_batched_all_pairs_uneven(cmp_engine):
    rng = np.random.Generator(np.random.PCG64(203))
    # digests of 1..6000 filters (16 KiB .. ~94 MiB at -b 16), cut from one
    # image so that pairs share chunks, plus some fresh ones
    img = synth.disk_image(204, 160 * MIB)
    counts = [1, 2, 7, 40, 130, 600, 1500, 6000] + [int(x) for x in rng.integers(1, 400, 40)]
    ds, bufs = [], []
    for k, c in enumerate(counts):
        n = c * 16 * KIB
        if k % 5 == 4:
            b = synth.segment_mix(rng, n)
        else:
            o = int(rng.integers(0, img.size - n + 1))
            b = img[o:o + n]
        bufs.append(b)
    names = [b"seg%03d" % k for k in range(len(bufs))]
    s = cmp_engine.new_set()
    assert np.all(s.digest(bufs, names, dd_block=16384) == 0)
    th = oracle.threads_available()
    ods = [oracle.digest_block(b, 16384, n, threads=th) for b, n in zip(bufs, names)]
    for i, od in enumerate(ods):
        same_digest(s, i, od)
    got = cmp_engine.compare_all(s)
    want = oracle.compare_all(ods, threads=th)
    assert np.array_equal(got, want), np.flatnonzero(got != want)[:10]
    assert (want >= 21).sum() > 20

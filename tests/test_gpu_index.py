"""The index paths on the GPU (sdhash --index / --index-search;
sdbf/sdbf_core.cc:110-115,157-177, sdhash-src/sdhash_threads.cc:376-492),
against the CPU oracle: every digest byte, the index bits and its element
count after a call, and every block's lookup counts."""
import numpy as np
import pytest

import oracle
import sdhash_amd
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


def indexed_inputs(seed, n, size):
    """Files that share content: each is cut from one image (so later ones
    repeat features of earlier ones) with fresh bytes mixed in."""
    rng = np.random.Generator(np.random.PCG64(seed))
    img = synth.disk_image(seed + 1, 4 * size)
    out = []
    for k in range(n):
        o = int(rng.integers(0, img.size - size))
        m = int(rng.integers(size // 4, size))
        out.append(np.concatenate([img[o:o + m], synth.segment_mix(rng, size - m)]))
    return out


@pytest.mark.parametrize("isize", [64 << 20, 64 << 10])
def test_index_stream_and_block_inputs_in_order(engine, isize):
    # stream inputs are gated by the index, block inputs insert into it; the
    # order given decides which features the gate skips
    files = indexed_inputs(401, 10, 700 * KIB)
    files.insert(4, synth.disk_image(402, 2 * MIB + 3000))       # block mode
    files.insert(7, synth.disk_image(403, 5 * 4096 + 600))       # block mode, dd 4096
    modes = [0] * len(files)
    modes[4], modes[7] = 16384, 4096
    names = [b"ix/%02d" % k for k in range(len(files))]
    ix = engine.new_index(isize)
    s = engine.new_set()
    st, _ = s.digest_indexed(files, names, dd_block=modes, index=ix)
    assert np.all(st == 0)
    oix = oracle.Index(isize)
    for k, (f, m) in enumerate(zip(files, modes)):
        od = oracle.digest_block_indexed(f, m, names[k], index=oix)[0] if m else \
            oracle.digest_stream_indexed(f, oix, names[k])
        same_digest(s, k, od)
    bits, ecount = ix.read()
    assert ecount == oix.elem_count, (ecount, oix.elem_count)
    assert np.array_equal(bits, oix.bf)
    # the gate changed some stream digest
    assert any(oracle.digest_stream(f).bf_count != s.info(k).bf_count
               for k, f in enumerate(files) if not modes[k])
    # a second call continues from the index's content
    more = indexed_inputs(404, 3, 300 * KIB)
    st, _ = s.digest_indexed(more, dd_block=0, index=ix)
    assert np.all(st == 0)
    for k, f in enumerate(more):
        same_digest(s, len(files) + k, oracle.digest_stream_indexed(f, oix))
    bits, ecount = ix.read()
    assert ecount == oix.elem_count and np.array_equal(bits, oix.bf)
    ix.close()


def test_index_block_insert_large(engine):
    # the --index hashing of large files: 16 KiB blocks, 64 MiB index
    img = synth.disk_image(405, 48 * MIB)
    parts = [img[:20 * MIB], img[12 * MIB:40 * MIB + 777], img[30 * MIB:]]
    ix = engine.new_index(64 << 20)
    s = engine.new_set()
    st, _ = s.digest_indexed(parts, dd_block=16384, index=ix)
    assert np.all(st == 0)
    oix = oracle.Index(64 << 20)
    for k, p in enumerate(parts):
        same_digest(s, k, oracle.digest_block_indexed(p, 16384, index=oix)[0])
    bits, ecount = ix.read()
    assert ecount == oix.elem_count and np.array_equal(bits, oix.bf)
    ix.close()


def test_index_search(engine):
    # --index-search: block inputs look up every fourth counted feature in
    # each index; the digests are the plain ones
    img = synth.disk_image(406, 24 * MIB)
    refs = []
    for k, (a, b) in enumerate(((0, 8 * MIB), (6 * MIB, 14 * MIB), (16 * MIB, 24 * MIB))):
        oix = oracle.Index(64 << 20 if k == 0 else 1 << 20)
        oracle.digest_block_indexed(img[a:b], 16384, index=oix)
        refs.append(oix)
    gix = []
    for r in refs:
        g = engine.new_index(r.size)
        g.write(r.bf, r.elem_count)
        gix.append(g)
    rng = np.random.Generator(np.random.PCG64(407))
    qs = [img[3 * MIB:7 * MIB], np.concatenate([img[15 * MIB:17 * MIB], synth.segment_mix(rng, MIB)]),
          synth.segment_mix(rng, 2 * MIB + 1000), img[:700]]
    modes = [16384, 4096, 16384, 16384]
    s = engine.new_set()
    st, got = s.digest_indexed(qs, dd_block=modes, search=gix)
    assert np.all(st == 0)
    want = []
    for k, (q, m) in enumerate(zip(qs, modes)):
        od, mm = oracle.digest_block_indexed(q, m, search=refs)
        same_digest(s, k, od)
        want.append(mm)
    want = np.concatenate(want)
    assert np.array_equal(got, want)
    assert (got > 4).sum() > 10 and (got[:, 0] > 4).any() and (got[:, 2] > 4).any()
    for g in gix:
        g.close()


def test_index_sizes_and_round_trip(engine):
    with pytest.raises(sdhash_amd.SdbfxError):
        engine.new_index(3000)
    with pytest.raises(sdhash_amd.SdbfxError):
        engine.new_index(32)
    ix = engine.new_index(4096)
    bits = np.random.Generator(np.random.PCG64(408)).integers(0, 256, 4096, dtype=np.uint8)
    ix.write(bits, 1234)
    b2, e2 = ix.read()
    assert e2 == 1234 and np.array_equal(b2, bits)
    ix.close()

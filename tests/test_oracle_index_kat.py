"""Known-answer tests of the oracle's index paths, every expected value
derived by hand from the reference's code (CPU only).

  info->index in stream mode: the gate between the filter insert and the big
      filter (sdbf/sdbf_core.cc:106-121)
  info->index in block mode: every counted feature inserted, the result
      ignored (sdbf/sdbf_core.cc:160-166)
  info->setlist in block mode: every fourth counted feature looked up in each
      set's index (sdbf/sdbf_core.cc:167-177, check_indexes at :482-493)
  bloom_filter: mask from the size, query_and_set (sdbf/bloom_filter.cc:52-71,
      334-359)

Features come from feature(t, h, top) of test_oracle_kat: filter positions
5t..5t+4, big-filter positions h << 11 | 5t+k, and with a 64 MiB index
(mask 2^29 - 1) index positions equal to the words themselves, so features
that differ in any of (t, h, top) share no index bit.
"""
import numpy as np
import pytest

import oracle
from test_oracle_kat import Seq, feature, filt


def test_index_mask_from_size():
    # BIT_MASKS_32[log_size + 1], log_size = bit length of the size
    assert oracle.Index(64).mask == 0x1FF
    assert oracle.Index(16384).mask == 0x1FFFF          # the 16 KiB big filter
    assert oracle.Index(64 << 20).mask == 0x1FFFFFFF    # the --index filter


def test_index_insert_and_query_by_brute_force():
    rng = np.random.default_rng(5)
    ix = oracle.Index(64)
    bits = np.zeros(512, bool)
    count = 0
    for _ in range(300):
        h = rng.integers(0, 2 ** 32, 5, dtype=np.uint64).astype(np.uint32)
        pos = [int(x) & 0x1FF for x in h]
        want_q = all(bits[p] for p in pos)
        assert ix.query(h) == want_q
        new = not want_q
        assert ix.insert(h) == new
        for p in pos:
            bits[p] = True
        count += new
        assert ix.elem_count == count
    assert np.array_equal(np.unpackbits(ix.bf, bitorder="little").astype(bool), bits)


def stream_indexed(seq, index):
    sc, w = seq.zeros(64).arrays()
    f, n, last = oracle.stream_hash_scored_indexed(sc, w, index)
    return f.reshape(n, 256), n, last


def test_stream_index_gate():
    ix = oracle.Index(64 << 20)
    assert ix.insert(feature(3, 1, top=1))
    s = Seq()
    s.add(feature(0, 0)).add(feature(1, 0))
    s.add(feature(3, 1, top=1))     # new filter bits (slot 3 set in filter 0), already in the
                                    # index: skipped (:110-115), never reaches the big filter
    s.add(feature(3, 2, top=2))     # slot 3 already set: 0 new bits, skipped before the index
    for t in range(4, 162):
        s.add(feature(t, 0))        # 158 more counted: 160, filter 0 full (:124-128)
    s.add(feature(5, 0, top=3))     # filter 1: new bits; new to the index (inserted, counted
                                    # there); big-filter bits those of feature(5, 0): rejected
    s.add(feature(3, 1, top=1))     # filter 1: new bits; in the index: skipped
    for t in range(6, 6 + 21):
        s.add(feature(t, 4))        # 21 counted in filter 1
    f, n, last = stream_indexed(s, ix)
    assert (n, last) == (2, 21)
    assert np.array_equal(f[0], filt([0, 1] + list(range(3, 162))))
    assert np.array_equal(f[1], filt([3, 5] + list(range(6, 27))))
    # the index: the pre-inserted feature, the 181 counted ones, feature(5, 0, top=3)
    assert ix.elem_count == 1 + 181 + 1
    assert ix.query(feature(5, 0, top=3))
    assert not ix.query(feature(3, 2, top=2))
    assert ix.query(feature(161, 0)) and ix.query(feature(6, 4))


def test_stream_index_shared_by_consecutive_inputs():
    # the second input's features that the first put in the index are skipped
    ix = oracle.Index(64 << 20)
    a = Seq()
    for t in range(30):
        a.add(feature(t, 0))
    f, n, last = stream_indexed(a, ix)
    assert (n, last) == (1, 30) and ix.elem_count == 30
    b = Seq()
    for t in range(40):
        b.add(feature(t, 0))        # 0..29 in the index: skipped (their filter bits kept)
    f, n, last = stream_indexed(b, ix)
    assert (n, last) == (1, 10) and ix.elem_count == 40
    assert np.array_equal(f[0], filt(range(40)))


def block_indexed(seq, threshold=16, allowed=0, index=None, search=()):
    sc, w = seq.zeros(64).arrays()
    return oracle.block_hash_scored_indexed(sc, w, threshold, allowed, index, search)


def test_block_index_insert():
    ix = oracle.Index(64 << 20)
    assert ix.insert(feature(2, 0, top=5))
    s = Seq()
    s.add(feature(0, 0))
    s.add(feature(0, 1))            # 0 new filter bits: not counted, not inserted
    s.add(feature(1, 0))
    s.add(feature(2, 0, top=5))     # counted; its index bits are all set: insert returns
                                    # false, the result is ignored (:163-166)
    s.add(feature(3, 0))
    bf, cnt, m = block_indexed(s, index=ix)
    assert cnt == 4
    assert np.array_equal(bf, filt(range(4)))
    assert ix.elem_count == 1 + 3
    assert ix.query(feature(0, 0)) and ix.query(feature(3, 0))
    assert not ix.query(feature(0, 1))


def test_block_index_search_every_fourth_counted_feature():
    s1, s2 = oracle.Index(64 << 20), oracle.Index(64 << 20)
    for x, t in ((s1, 0), (s1, 4), (s1, 9), (s2, 1), (s2, 8)):
        x.insert(feature(t, 0))
    ins = oracle.Index(64 << 20)
    s = Seq()
    for t in range(4):
        s.add(feature(t, 0))        # counted 0..3: feature 0 looked up (count 0)
    s.add(feature(2, 7))            # 0 new bits: not counted, no lookup
    for t in range(4, 12):
        s.add(feature(t, 0))        # counted 4..11: features 4 and 8 looked up
    bf, cnt, m = block_indexed(s, index=ins, search=(s1, s2))
    assert cnt == 12
    # s1 holds features 0, 4 (looked up) and 9 (not); s2 holds 1 (not) and 8
    assert list(m) == [2, 1]
    # with searched indexes the insert index is left alone (:162-166)
    assert ins.elem_count == 0 and not ins.bf.any()


def test_block_search_lookup_needs_all_five_bits():
    s1 = oracle.Index(64 << 20)
    w = feature(0, 0)
    s1.insert(w[:4] + [w[4] ^ (1 << 20)])   # four of the five bits
    s = Seq().add(feature(0, 0))
    for t in range(1, 5):
        s.add(feature(t, 0))
    bf, cnt, m = block_indexed(s, search=(s1,))
    assert cnt == 5 and list(m) == [0]


@pytest.mark.parametrize("size", [64 << 20])
def test_digest_entries_agree_with_the_scored_paths_on_real_data(size):
    # the byte-level entries take the same gate: a second digest of the same
    # bytes into the same index finds every counted feature already there
    from sdhash_amd import synth
    data = synth.file_set(71, 1, 200 * 1024)[0]
    ix = oracle.Index(size)
    d1 = oracle.digest_stream_indexed(data, ix)
    plain = oracle.digest_stream(data)
    assert d1.bf_count == plain.bf_count and np.array_equal(d1.filters, plain.filters)
    n1 = ix.elem_count
    assert n1 == (d1.bf_count - 1) * 160 + d1.last_count or d1.last_count == 160
    d2 = oracle.digest_stream_indexed(data, ix)
    assert ix.elem_count == n1 and d2.bf_count == 1 and d2.last_count == 0
    # block mode: inserting the digest's features leaves a search of the
    # same bytes finding every looked-up feature
    bx = oracle.Index(size)
    db, _ = oracle.digest_block_indexed(data, 16384, index=bx)
    _, m = oracle.digest_block_indexed(data, 16384, search=(bx,))
    want = [(int(c) + 3) // 4 for c in db.elem_counts]
    assert list(m[:, 0]) == want

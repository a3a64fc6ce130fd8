"""Known-answer tests of the oracle's hashing and scoring rules, with every
expected value derived by hand from the reference's code (CPU only).

The oracle's scored entry points take per-position popularity scores and
SHA-1 words directly, so each feature's bit positions are chosen here:
feature(slot t, tag h) has the five 256-byte-filter positions 5t..5t+4
(w & 0x7FF) and the 16 KiB de-duplication-filter positions h << 11 | 5t+k
(w & 0x1FFFF).  Features with different (t, h) never share a bit of the big
filter; features with the same t share their five filter bits.

  stream: gen_chunk_hash (sdbf/sdbf_core.cc:94-136), the 32 MiB unit loop and
          the final trim (sdbf/sdbf_core.cc:197-246)
  block:  gen_block_hash (sdbf/sdbf_core.cc:142-191), the threshold walk and
          the tail rule (sdbf/sdbf_core.cc:252-324)
  score:  sdbf_score / sdbf_max_score (sdbf/sdbf_core.cc:331-455)
"""
import numpy as np
import pytest

import oracle


def feature(t, h=0, top=0):
    return [(top << 17) | ((h & 63) << 11) | (5 * t + k) for k in range(5)]


def filt(slots):
    """A 256-byte filter with the five bits of each slot set."""
    v = np.zeros(2048, np.uint8)
    for t in slots:
        v[5 * t:5 * t + 5] = 1
    return np.packbits(v, bitorder="little")


class Seq:
    """Per-position scores and SHA-1 words; positions without a feature
    score 0."""

    def __init__(self):
        self.sc, self.w = [], []

    def add(self, words, score=17):
        self.sc.append(score)
        self.w.append(words)
        return self

    def zeros(self, n):
        for _ in range(n):
            self.add([0] * 5, 0)
        return self

    def arrays(self):
        return np.array(self.sc, np.uint16), np.array(self.w, np.uint32).reshape(-1, 5)


def stream(seq, unit=32 << 20):
    # a unit's last 64 positions are never scanned (sdbf_core.cc:103)
    sc, w = seq.zeros(64).arrays()
    f, n, last = oracle.stream_hash_scored(sc, w, unit)
    return f.reshape(n, 256), n, last


# ---------------------------------------------------------------- stream mode

def test_stream_zero_new_bit_feature_is_skipped_before_the_big_filter():
    s = Seq().add(feature(0, 0))
    s.add(feature(0, 5))            # bits 0..4 already set in filter 0: 0 new, skipped (:108-109)
    for t in range(1, 160):
        s.add(feature(t, 0))        # 160 counted: filter 0 is full (:124-128)
    s.add(feature(0, 5))            # fresh filter 1: new bits; its big-filter bits are unset,
                                    # because the skipped copy never reached the big filter: counted
    s.add(feature(150, 9), score=16)  # score 16 is not above the threshold (:102): ignored
    for t in range(1, 21):
        s.add(feature(t, 6))
    f, n, last = stream(s)
    assert (n, last) == (2, 21)
    assert np.array_equal(f[0], filt(range(160)))
    assert np.array_equal(f[1], filt(range(21)))


def test_stream_big_filter_rejection_keeps_the_filter_bits():
    s = Seq()
    for t in range(160):
        s.add(feature(t, 0))
    s.add(feature(0, 0))            # new in filter 1, all five bits set in the big filter:
                                    # rejected (:119-121) -- its bits stay in filter 1 (:108)
    s.add(feature(0, 7))            # so this one has 0 new bits in filter 1: skipped
    for t in range(1, 26):
        s.add(feature(t, 8))
    f, n, last = stream(s)
    assert (n, last) == (2, 25)
    assert np.array_equal(f[1], filt(range(26)))


@pytest.mark.parametrize("count,want", [
    (10, (1, 10)),
    (159, (1, 159)),
    (160, (1, 160)),   # rollover at 160 -> (2, 0), then the trim: 0 < 20 -> (1, 160)
    (179, (1, 160)),   # (2, 19) trimmed
    (180, (2, 20)),    # 20 is not below max_elem / 8: kept
    (200, (2, 40)),
    (320, (2, 160)),   # (3, 0) trimmed
])
def test_stream_rollover_at_160_and_trim(count, want):
    s = Seq()
    for t in range(count):
        s.add(feature(t % 160, t // 160))
    f, n, last = stream(s)
    assert (n, last) == want
    full = count // 160
    for i in range(n):
        if i < full:
            assert np.array_equal(f[i], filt(range(160)))
        else:
            assert np.array_equal(f[i], filt(range(count % 160)))


def test_stream_big_filter_restarts_after_8738_features():
    s = Seq()
    for t in range(8738):               # 54 full filters + 98 in filter 54
        s.add(feature(t % 160, t // 160))
    s.add(feature(98, 1))               # the words of feature 258 (counted long ago), at a free
                                        # slot of filter 54: the 8738th feature started a new,
                                        # empty big filter (:129-132), so this one is counted
    f, n, last = stream(s)
    assert (n, last) == (55, 99)


def test_stream_big_filter_not_restarted_at_8737():
    s = Seq()
    for t in range(8737):               # filter 54 holds slots 0..96
        s.add(feature(t % 160, t // 160))
    s.add(feature(97, 1))               # words of feature 257: all five bits in the big filter
    s.add(feature(98, 54))              # new: the 8738th counted (restart after it)
    f, n, last = stream(s)
    assert (n, last) == (55, 98)
    assert np.array_equal(f[54], filt(range(99)))   # the rejected feature's bits stay


def test_stream_restart_count_is_per_unit():
    # units of 5064 positions scan 5000 each; bigfi_count starts at 0 in every
    # call of gen_chunk_hash (:99), so 5000 + 5000 features never restart the
    # big filter
    unit = 5064
    s = Seq()
    for t in range(5000):
        s.add(feature(t % 160, t // 160))
    s.zeros(64)
    for t in range(5000, 10000):
        s.add(feature(t % 160, t // 160))
    s.zeros(64)
    # the remainder unit: 62 full filters + 80 in filter 62 so far
    s.add(feature(80, 1))               # words of feature 240: still in the big filter: rejected
    for j in range(30):
        s.add(feature(81 + j, 62))
    s.zeros(1000 - 31 - 64)
    f, n, last = stream(s, unit)
    assert (n, last) == (63, 110)


def test_stream_big_filter_persists_across_units():
    # the big filter is the file's (big_filters->back()), not the unit's
    unit = 264
    s = Seq()
    for t in range(200):                # unit 1 scans 200 positions
        s.add(feature(t % 160, t // 160))
    s.zeros(64)
    s.add(feature(40, 0))               # words of feature 40, at free slot 40 of filter 1: rejected
    for j in range(30):
        s.add(feature(41 + j, 1))
    s.zeros(264 - 31 - 64)
    f, n, last = stream(s, unit)
    assert (n, last) == (2, 70)


# ---------------------------------------------------------------- block mode

def block_hash(seq, threshold, allowed):
    sc, w = seq.zeros(64).arrays()
    return oracle.block_hash_scored(sc, w, threshold, allowed)


def test_block_allowance_is_spent_only_on_counted_features():
    s = Seq()
    s.add(feature(0), 20)
    s.add(feature(0, 3), 20)   # 0 new bits: skipped without count (:161-162); allowance kept
    s.add(feature(1), 20)      # counted: allowance 2 -> 0 (:178-179)
    s.add(feature(2), 20)      # score == threshold with no allowance: not selected (:157-158)
    s.add(feature(3), 21)      # above the threshold: selected whatever the allowance
    s.add(feature(4), 19)      # below: never
    bf, cnt = block_hash(s, 20, 2)
    assert cnt == 3
    assert np.array_equal(bf, filt([0, 1, 3]))
    s = Seq().add(feature(0), 20).add(feature(3), 21)
    bf, cnt = block_hash(s, 20, 0)
    assert cnt == 1 and np.array_equal(bf, filt([3]))


def test_block_hashing_stops_at_192():
    s = Seq()
    for t in range(200):
        s.add(feature(t), 30)
        if t == 10:
            s.add(feature(3, 9), 30)   # a duplicate: not counted
    bf, cnt = block_hash(s, 20, 0)
    assert cnt == 192
    assert np.array_equal(bf, filt(range(192)))


def block_digest(seq, dd):
    sc, w = seq.arrays()
    f, ec = oracle.block_digest_scored(sc, w, dd)
    return f.reshape(-1, 256), ec


def test_block_threshold_and_allowance_from_histogram():
    # dd 1024: 960 scanned positions.  150 features of score 20, then 100 of
    # score 30: the walk from k = 65 down has sum 100 at k = 20, where
    # 100 + 150 > 192 stops it: threshold 20, allowance 92 (:271-276).  In
    # position order the first 92 score-20 features are counted, the other
    # 58 are not selected, and all 100 of score 30 are counted.
    s = Seq()
    for t in range(150):
        s.add(feature(t), 20)
    for t in range(150, 250):
        s.add(feature(t), 30)
    s.zeros(1024 - 250)
    f, ec = block_digest(s, 1024)
    assert list(ec) == [192]
    assert np.array_equal(f[0], filt(list(range(92)) + list(range(150, 250))))


def test_block_walk_exits_at_k_15():
    # 10 features of score 17 never take the sum past 192: the walk ends with
    # k = 15 and allowance 182, so score-15 features are selected too
    s = Seq()
    for t in range(10):
        s.add(feature(t), 17)
    for t in range(10, 15):
        s.add(feature(t), 15)
    s.add(feature(20), 14)
    s.zeros(1024 - 16)
    f, ec = block_digest(s, 1024)
    assert list(ec) == [15]
    assert np.array_equal(f[0], filt(range(15)))


def test_block_tail_uses_threshold_16_and_allowance_192():
    # a 600-byte tail (>= 512 is kept, :309) is hashed with threshold 16 and
    # allowance 192 (:317): scores 16 and 17 are selected, 15 is not.  The
    # histogram walk would have given k = 15 and selected all three.
    s = Seq().zeros(1024)
    s.add(feature(0), 16).add(feature(1), 15).add(feature(2), 17)
    s.zeros(600 - 3)
    f, ec = block_digest(s, 1024)
    assert list(ec) == [0, 2]
    assert np.array_equal(f[1], filt([0, 2]))
    # a 500-byte tail is dropped
    s = Seq().zeros(1024 + 500)
    f, ec = block_digest(s, 1024)
    assert len(ec) == 1


# ---------------------------------------------------------------- scores

def bits(*ranges):
    v = np.zeros(2048, np.uint8)
    for lo, hi in ranges:
        v[lo:hi] = 1
    return np.packbits(v, bitorder="little")


X = bits((0, 600))            # 600 bits
Z = bits((1200, 1800))        # 600 bits, disjoint from X


def dd_digest(name, filters, ecs):
    return oracle.Digest(name, np.concatenate(filters), len(filters), 0, np.array(ecs, np.uint16), 100000,
                         16384)


def stream_digest(name, filters, last):
    return oracle.Digest(name, np.concatenate(filters), len(filters), last, None, 100000, 0)


def test_score_fewer_filters_is_the_reference():
    # a (1 filter) is the reference side whichever argument it is (:349-359):
    # X matches X fully -> 100.  With b as the reference: (1 + 0) / 2 -> 50.
    a = stream_digest(b"a", [X], 160)
    b = stream_digest(b"b", [X, Z], 160)
    assert oracle.score(a, b) == 100
    assert oracle.score(b, a) == 100


def test_score_equal_counts_swap_on_last_count_or_name():
    # equal filter counts: swap when the first digest's last element count is
    # larger OR its name sorts after (:351-353).  a = [X, X] (last count 180),
    # b = [X, Z] (last count 170).
    a = dd_digest(b"a", [X, X], [160, 180])
    b = dd_digest(b"b", [X, Z], [160, 170])
    # (a, b): 180 > 170 -> swap -> reference b: X -> 1, Z -> 0 -> 50
    assert oracle.score(a, b) == 50
    # (b, a): 170 > 180 fails, but "b" > "a" -> swap -> reference a: 100
    assert oracle.score(b, a) == 100
    # equal last counts: the name alone decides
    c = dd_digest(b"c", [X, Z], [160, 180])
    assert oracle.score(a, c) == 100     # 180 > 180 fails, "a" < "c": reference a
    assert oracle.score(c, a) == 100     # "c" > "a": swap -> reference a


def test_score_minus_one_sentinel_depends_on_filter_order():
    # Every filter of t is sparse (< 16 elements), so a non-sparse reference
    # filter scores -1 (no target filter qualifies, :413-414) and a sparse one
    # 0 (:404-405).  The running sum starts at -1 and restarts from the next
    # value while it is negative (:379): [0, -1, -1] -> -1;
    # [-1, -1, 0] -> 0 (denominator 3 - 1 sparse = 2, :382-385);
    # [-1, 0, -1] -> -1.
    t = dd_digest(b"t", [X, X, X, X], [5, 5, 5, 5])
    assert oracle.score(dd_digest(b"a", [X, X, X], [5, 100, 100]), t) == -1
    assert oracle.score(dd_digest(b"a", [X, X, X], [100, 100, 5]), t) == 0
    assert oracle.score(dd_digest(b"a", [X, X, X], [100, 5, 100]), t) == -1


def test_score_denominator_single_filter():
    # one sparse reference filter: 0 / 1 -> 0, because the sparse filters are
    # taken off the denominator only when there is more than one (:383-384)
    t = dd_digest(b"t", [X, X], [160, 160])
    assert oracle.score(dd_digest(b"a", [X], [5]), t) == 0
    # two sparse reference filters: denominator 0 -> -1 (:386-387)
    t3 = dd_digest(b"t", [X, X, X], [160, 160, 160])
    assert oracle.score(dd_digest(b"a", [X, X], [5, 5]), t3) == -1


def test_score_rounds_half_up():
    # s1 = s2 = 160 -> cutoff CUTOFFS256[4096 / 320 = 12] = 240.  x has 1040
    # bits; y has 340 of them and 700 others, so max_est = 1040 and the match
    # is 340: (340 - 240) / (1040 - 240) = 0.125 -> 12.5 -> boost round = 13.
    x = bits((0, 1040))
    y = bits((0, 340), (1100, 1800))
    assert oracle.score(stream_digest(b"a", [x], 160), stream_digest(b"b", [y], 160)) == 13
    # one bit less: 99 / 800 -> 12.375 -> 12
    y2 = bits((0, 339), (1100, 1801))
    assert oracle.score(stream_digest(b"a", [x], 160), stream_digest(b"b", [y2], 160)) == 12


def test_score_sampling_takes_the_first_filters():
    # sample 2 (sdbf_core.cc:342-347): a = [X, X, Z] counts as its first two
    # filters, and the swap rule reads the element count of filter 1, which
    # for a stream digest is max_elem (160), not a's last count (100).
    W = bits((1900, 2048))
    a = stream_digest(b"a", [X, X, Z], 100)
    b_lo = stream_digest(b"b", [X, W], 150)   # 160 > 150: swap -> reference b: (1 + 0) / 2
    b_hi = stream_digest(b"b", [X, W], 170)   # 160 > 170 fails, "a" < "b": reference a's X, X -> 100
    assert oracle.score(a, b_lo, 2) == 50
    assert oracle.score(a, b_hi, 2) == 100
    # sample 1: a counts as one filter (< 2): reference a's first filter X -> 100
    assert oracle.score(a, b_lo, 1) == 100
    # no sampling, or a sample of at least a's count: a has 3 > 2 filters ->
    # reference b -> 50
    assert oracle.score(a, b_hi, 0) == 50
    assert oracle.score(a, b_hi, 3) == 50

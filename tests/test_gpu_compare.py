"""Comparison parity on the GPU: every 0-100 score (and -1) identical with
the CPU oracle, including the reference's orientation rule, for both
comparison kernels (VALU popcount, MFMA on FP4 bit elements and MFMA on
int8 byte planes)."""
import numpy as np
import pytest

import oracle
import sdhash_amd
from sdhash_amd import synth

pytestmark = pytest.mark.gpu

KIB = 1024
MIB = 1 << 20


def upload_oracle(engine, ds):
    s = engine.new_set()
    for d in ds:
        s.add(d.name, d.orig_size, d.bf_count, d.last_count, d.dd_block_size, d.max_elem, d.filters,
              d.elem_counts)
    return s


def test_compare_all_stream(cmp_engine):
    files = synth.file_set(31, 24, 256 * KIB)
    names = [b"f%02d" % i for i in range(len(files))]
    ods = [oracle.digest_stream(f, n) for f, n in zip(files, names)]
    want = oracle.compare_all(ods)
    s = cmp_engine.new_set()
    assert np.all(s.digest(files, names) == 0)
    got = cmp_engine.compare_all(s)
    assert np.array_equal(got, want), np.flatnonzero(got != want)[:10]
    assert (want > 0).sum() > 10


def test_compare_all_mixed_modes_and_names(cmp_engine):
    files = synth.file_set(32, 10, 200 * KIB)
    # names chosen so that the strcmp part of the orientation rule matters
    names = [b"z", b"a", b"m", b"b", b"y", b"c", b"x", b"d", b"w", b"e"]
    ods = []
    for i, (f, n) in enumerate(zip(files, names)):
        ods.append(oracle.digest_block(f, 16384, n) if i % 3 == 0 else oracle.digest_stream(f, n))
    s = upload_oracle(cmp_engine, ods)
    want = oracle.compare_all(ods)
    got = cmp_engine.compare_all(s)
    assert np.array_equal(got, want)


def test_compare_to_fragments_vs_image(cmp_engine):
    img = synth.disk_image(33, 2 * MIB)
    frags = synth.fragments(34, img, 256)
    tgt = [oracle.digest_block(img[:MIB], 16384, b"img.0000M"),
           oracle.digest_block(img[MIB:], 16384, b"img.0001M")]
    q = [oracle.digest_stream(f, b"frag%04d" % i) for i, f in enumerate(frags)]
    want = oracle.compare_to(q, tgt)
    gq = cmp_engine.new_set()
    assert np.all(gq.digest(frags, [d.name for d in q]) == 0)
    gt = cmp_engine.new_set()
    assert np.all(gt.digest([img[:MIB], img[MIB:]], [b"img.0000M", b"img.0001M"], dd_block=16384) == 0)
    got = cmp_engine.compare_to(gq, gt)
    assert np.array_equal(got, want)
    assert (want[:128] > 0).sum() > 50


def test_compare_hand_cases(cmp_engine):
    def bits(*ranges):
        v = np.zeros(2048, np.uint8)
        for lo, hi in ranges:
            v[lo:hi] = 1
        return np.packbits(v, bitorder="little")
    a = oracle.Digest(b"a", bits((0, 600)), 1, 160, None, 1000, 0)
    b = oracle.Digest(b"b", bits((0, 500), (1500, 1600)), 1, 160, None, 1000, 0)
    c = oracle.Digest(b"c", bits((0, 90), (1024, 1624)), 1, 160, None, 1000, 0)
    d = oracle.Digest(b"d", bits((0, 90), (1024, 1624)), 1, 160, None, 1000, 0)
    e = oracle.Digest(b"e", np.concatenate([bits((0, 600)), bits((0, 600))]), 2, 0,
                      np.array([5, 5], np.uint16), 1000, 16384)
    ds = [a, b, c, d, e]
    s = upload_oracle(cmp_engine, ds)
    got = cmp_engine.compare_all(s)
    want = oracle.compare_all(ds)
    assert np.array_equal(got, want)
    assert got[0] == 72          # a|b
    assert got[3] == -1          # a|e: every target filter sparse


def test_compare_extreme_bit_patterns(cmp_engine):
    # Filters whose bits sit in single operand registers of the MFMA forms
    # (only bit 0, only bit 3 of every nibble; byte planes 0 and 7), full
    # and empty words at the ends of the K steps, all 2048 bits set (the
    # largest accumulator sum), disjoint halves (counts of 0) and prefixes
    # that pass or fail the early exits of bf_bitcount_cut_256.
    def rep(b):
        return np.full(256, b, np.uint8)
    def words(*idx):
        v = np.zeros(256, np.uint8)
        for i in idx:
            v[4 * i:4 * i + 4] = 0xFF
        return v
    def prefix(n):
        v = np.zeros(256, np.uint8)
        v[:n] = 0xFF
        return v
    pats = [rep(0xFF), rep(0x55), rep(0xAA), rep(0x11), rep(0x88), rep(0x01), rep(0x80),
            words(0, 63), words(*range(0, 64, 2)), words(*range(1, 64, 2)), prefix(32), prefix(64),
            prefix(128), prefix(160), np.concatenate([rep(0xFF)[:128], rep(0x0F)[:128]])]
    groups = [(0, 1, 2, 3), (4, 5, 6, 7, 8), (9, 10, 11, 12), (13, 14, 0), (1, 4, 9, 13), (2, 5, 10, 14), (0,)]
    ds = []
    for k, g in enumerate(groups):
        f = np.concatenate([pats[i] for i in g])
        ds.append(oracle.Digest(b"x%d" % k, f, len(g), 160, None, 100000, 0))
    s = upload_oracle(cmp_engine, ds)
    want = oracle.compare_all(ds)
    got = cmp_engine.compare_all(s)
    assert np.array_equal(got, want), (got, want)
    assert (want == 100).any() and (want > 0).sum() >= 5
    # and one against many: the all-ones filter facing every pattern
    q = [oracle.Digest(b"q", rep(0xFF), 1, 160, None, 1000, 0)]
    t = [oracle.Digest(b"t%02d" % i, p, 1, 160, None, 1000, 0) for i, p in enumerate(pats)]
    want = oracle.compare_to(q, t)
    got = cmp_engine.compare_to(upload_oracle(cmp_engine, q), upload_oracle(cmp_engine, t))
    assert np.array_equal(got, want), (got, want)


def test_compare_large_target_chunks(cmp_engine):
    # a target digest of > 1024 filters is split over several work items
    img = synth.disk_image(35, 20 * MIB)
    frags = synth.fragments(36, img, 64, size=64 * KIB)
    t = [oracle.digest_block(img, 16384, b"img")]
    q = [oracle.digest_stream(f, b"q%02d" % i) for i, f in enumerate(frags)]
    want = oracle.compare_to(q, t)
    gq = upload_oracle(cmp_engine, q)
    gt = upload_oracle(cmp_engine, t)
    got = cmp_engine.compare_to(gq, gt)
    assert np.array_equal(got, want)
    # and the transposed call exercises the other orientation
    assert np.array_equal(cmp_engine.compare_to(gt, gq), oracle.compare_to(t, q))


def test_compare_identical_digests(cmp_engine):
    # every filter of every digest has exact duplicates: the MFMA path's
    # candidate queue sees a match in every row of every tile
    f = synth.file_set(37, 1, 2 * MIB)[0]
    names = [b"dup%02d" % i for i in range(12)]
    d = oracle.digest_stream(f, b"x")
    ods = [oracle.Digest(n, d.filters, d.bf_count, d.last_count, d.elem_counts, d.orig_size,
                         d.dd_block_size) for n in names]
    s = upload_oracle(cmp_engine, ods)
    got = cmp_engine.compare_all(s)
    want = oracle.compare_all(ods)
    assert np.array_equal(got, want)
    assert np.all(want == 100)


def test_compare_to_many_rows(cmp_engine):
    # several 128-instance row blocks and 1024-filter chunks in one group,
    # with reference rows of both sets (orientation flips on size)
    img = synth.disk_image(38, 24 * MIB)
    frags = synth.fragments(39, img, 600, size=8 * KIB)
    frags += [img[o:o + 4 * MIB] for o in (0, 9 * MIB)]
    t = [oracle.digest_block(img[:16 * MIB], 16384, b"img.0000M"),
         oracle.digest_block(img[16 * MIB:], 16384, b"img.0016M"),
         oracle.digest_block(img[9 * MIB:10 * MIB], 16384, b"img.small")]
    q = [oracle.digest_stream(x, b"q%03d" % i) for i, x in enumerate(frags)]
    q = [d for d in q if d is not None]
    want = oracle.compare_to(q, t)
    got = cmp_engine.compare_to(upload_oracle(cmp_engine, q), upload_oracle(cmp_engine, t))
    assert np.array_equal(got, want), np.argwhere(got != want)[:10]
    assert (want > 0).sum() > 100


def test_compare_repeated_filter(cmp_engine):
    # one filter repeated in two digests: every pair of a 128 x 128 tile
    # qualifies (the MFMA kernel then re-checks the whole tile exactly)
    d = oracle.digest_stream(synth.file_set(40, 1, 256 * KIB)[0], b"x")
    one = d.filters[:256]
    reps = [oracle.Digest(b"r%d" % i, np.tile(one, n), n, 160, None, n * 40000, 0)
            for i, n in enumerate((300, 260, 129))]
    s = upload_oracle(cmp_engine, reps)
    want = oracle.compare_all(reps)
    assert np.array_equal(cmp_engine.compare_all(s), want)
    assert np.all(want == 100)


def test_compare_pruned_multi_chunk(cmp_engine):
    # Enough row blocks and target filters that the MFMA kernel takes
    # 2048-filter chunks (16 tiles per work item, running maxima shared
    # across the 10 chunks of each row block).  One target digest has one
    # element count everywhere (uniform tiles), the other counts of 8-192
    # (mixed tiles, sparse filters, and low cutoffs that make most pairs
    # candidates).  Query filters are thinned copies of target filters plus
    # noise, so the best fraction of a row varies widely.
    rng = np.random.Generator(np.random.PCG64(41))
    nt, nq, per = 20000, 30, 300

    def bits(n, tenths):
        return rng.integers(0, 10, (n, 2048), dtype=np.uint8) < tenths

    tb = [bits(nt, 3), bits(nt, 3)]
    tec = [np.full(nt, 192, np.uint16), rng.integers(8, 193, nt).astype(np.uint16)]
    t = [oracle.Digest(b"t%d" % k, np.packbits(tb[k], axis=1, bitorder="little").reshape(-1), nt, 0, tec[k],
                       nt * 16384, 16384) for k in range(2)]
    q = []
    for d in range(nq):
        src, which = rng.integers(0, nt, per), rng.integers(0, 2, per)
        b = np.where(which[:, None] == 0, tb[0][src], tb[1][src])
        keep = rng.uniform(0.2, 1.0, per)[:, None]
        b = (b & (rng.random((per, 2048), dtype=np.float32) < keep)) | bits(per, 0.5)
        q.append(oracle.Digest(b"q%02d" % d, np.packbits(b, axis=1, bitorder="little").reshape(-1), per,
                               100 + d, None, per * 12000, 0))
    want = oracle.compare_to(q, t, threads=oracle.threads_available())
    got = cmp_engine.compare_to(upload_oracle(cmp_engine, q), upload_oracle(cmp_engine, t))
    assert np.array_equal(got, want), np.argwhere(got != want)[:10]
    assert len(set(want.reshape(-1).tolist())) > 10


def test_compare_target_set_grows(cmp_engine):
    # the MFMA kernel orders each target digest's filters by size code on
    # first use (col_order_kernel); digests appended to the set later are
    # ordered on their own and the earlier order is carried over
    img = synth.disk_image(91, 40 * MIB)
    frags = synth.fragments(92, img, 300, 8 * KIB)
    th = oracle.threads_available()
    t1 = oracle.digest_block(img[:24 * MIB], 16384, b"a", threads=th)
    t2 = oracle.digest_block(img[24 * MIB:], 4096, b"b", threads=th)
    q = [oracle.digest_stream(f, b"q%03d" % i) for i, f in enumerate(frags)]
    gq = upload_oracle(cmp_engine, q)
    gt = upload_oracle(cmp_engine, [t1])
    got = cmp_engine.compare_to(gq, gt)
    want = oracle.compare_to(q, [t1], threads=th)
    assert np.array_equal(got, want), np.argwhere(got != want)[:10]
    gt.add(t2.name, t2.orig_size, t2.bf_count, t2.last_count, t2.dd_block_size, t2.max_elem, t2.filters,
           t2.elem_counts)
    got = cmp_engine.compare_to(gq, gt)
    want = oracle.compare_to(q, [t1, t2], threads=th)
    assert np.array_equal(got, want), np.argwhere(got != want)[:10]
    assert (want > 0).sum() > 100


def test_compare_fast_mode(cmp_engine):
    # sdhash --fast: every filter folded to 64 bytes on the device; the folded
    # filters, Hamming weights and all scores equal the oracle's sdbf::fast
    files = synth.file_set(93, 16, 256 * KIB)
    img = synth.disk_image(94, 6 * MIB)
    names = [b"f%02d" % i for i in range(len(files))] + [b"img"]
    modes = [16384 if i % 4 == 0 else 0 for i in range(len(files))] + [16384]
    bufs = files + [img]
    ods = [oracle.digest_block(b, m, n) if m else oracle.digest_stream(b, n) for b, n, m in zip(bufs, names, modes)]
    s = cmp_engine.new_set()
    assert np.all(s.digest(bufs, names, dd_block=modes) == 0)
    s.fold()
    fods = [d.folded() for d in ods]
    for i, d in enumerate(fods):
        f, ec, hm = s.read(i)
        assert np.array_equal(f, d.filters) and np.array_equal(hm, d.hamming[: d.bf_count]), i
    got = cmp_engine.compare_all(s)
    want = oracle.compare_all(fods)
    assert np.array_equal(got, want), np.flatnonzero(got != want)[:10]
    assert (want > 0).sum() > 10
    # sdhash -c q t --fast: only the query set is folded (sdbf_set.cc:367-376);
    # each pair is scored by its reference digest's mode, so fragments (the
    # reference side) score folded against the plain image
    frags = synth.fragments(95, img, 200, 8 * KIB)
    q = [oracle.digest_stream(f, b"q%03d" % i) for i, f in enumerate(frags)]
    gq = upload_oracle(cmp_engine, q)
    gt = upload_oracle(cmp_engine, [ods[-1]])
    gq.fold()
    got = cmp_engine.compare_to(gq, gt)
    want = oracle.compare_to([d.folded() for d in q], [ods[-1]])
    assert np.array_equal(got, want), np.argwhere(got != want)[:10]
    # (folded 64 bytes against a plain filter's first 64 bytes: nearly
    # every score is 0, as in the reference)
    # both folded
    gt.fold()
    got = cmp_engine.compare_to(gq, gt)
    want = oracle.compare_to([d.folded() for d in q], [fods[-1]])
    assert np.array_equal(got, want), np.argwhere(got != want)[:10]
    assert (want[:100] > 0).sum() > 10   # folding loses matches: 25 of 100 here


def test_compare_folded_target_plain_reference(cmp_engine):
    # a plain reference digest against a folded target: 256-byte rules with the
    # target's folded Hamming weight (sdbf_core.cc:409-441), unsigned
    # denominators, fractions above 1 (scores above 100), and the pair whose
    # denominator is 0 (the reference aborts on it) refused
    img = synth.disk_image(96, 3 * MIB)
    big = [oracle.digest_block(img[o:o + MIB], 16384, b"b%d" % k) for k, o in enumerate((0, MIB, 2 * MIB))]
    small = [oracle.digest_stream(img[o:o + 40 * KIB], b"s%d" % k) for k, o in enumerate((5000, 1500000, 2600000))]
    gb, gs = upload_oracle(cmp_engine, big), upload_oracle(cmp_engine, small)
    gb.fold()   # the larger digests folded: the small plain ones are the reference side
    got = cmp_engine.compare_to(gs, gb)
    want = oracle.compare_to(small, [d.folded() for d in big])
    assert np.array_equal(got, want), (got, want)
    assert want.max() > 100

    def bits(*ranges):
        v = np.zeros(2048, np.uint8)
        for lo, hi in ranges:
            v[lo:hi] = 1
        return np.packbits(v, bitorder="little")
    # the target filter folds to 240 bits = the cutoff for 160 + 160 elements
    f = bits((0, 240), (512, 752), (1024, 1264), (1536, 1656))
    ref = oracle.Digest(b"r", f, 1, 160, None, 1, 0)
    tgt = oracle.Digest(b"t", np.concatenate([f, f]), 2, 160, None, 1, 0)
    assert oracle.score(ref, tgt.folded()) == -(1 << 31)
    g1, g2 = upload_oracle(cmp_engine, [ref]), upload_oracle(cmp_engine, [tgt])
    g2.fold()
    with pytest.raises(sdhash_amd.SdbfxError, match="not finite"):
        cmp_engine.compare_to(g1, g2)


def test_compare_sampled(cmp_engine):
    # sdhash -s N: a query digest of more than N filters takes part with its
    # first N (sdbf_core.cc:342-347)
    img = synth.disk_image(97, 8 * MIB)
    t = [oracle.digest_block(img[:4 * MIB], 16384, b"t0"), oracle.digest_stream(img[4 * MIB:5 * MIB], b"t1")]
    q = [oracle.digest_block(img[o:o + 600 * KIB], 16384, b"q%d" % k) for k, o in enumerate(range(0, 8 * MIB, MIB))]
    q += [oracle.digest_stream(img[o:o + 300 * KIB], b"r%d" % k) for k, o in enumerate(range(0, 8 * MIB, MIB))]
    gq, gt = upload_oracle(cmp_engine, q), upload_oracle(cmp_engine, t)
    for smp in (1, 5, 20, 64):
        want = oracle.compare_to(q, t, sample=smp)
        assert np.array_equal(cmp_engine.compare_to(gq, gt, sample=smp), want), smp
        assert np.array_equal(cmp_engine.compare_to(gt, gq, sample=smp), oracle.compare_to(t, q, sample=smp)), smp
    assert not np.array_equal(oracle.compare_to(q, t, sample=5), oracle.compare_to(q, t))


def test_compare_hand_derived_rules(cmp_engine):
    # the known answers of the oracle's score tests, on the device
    def bits(*ranges):
        v = np.zeros(2048, np.uint8)
        for lo, hi in ranges:
            v[lo:hi] = 1
        return np.packbits(v, bitorder="little")
    X, Z = bits((0, 600)), bits((1200, 1800))

    def dd(name, fs, ecs):
        return oracle.Digest(name, np.concatenate(fs), len(fs), 0, np.array(ecs, np.uint16), 100000, 16384)

    def st(name, fs, last):
        return oracle.Digest(name, np.concatenate(fs), len(fs), last, None, 100000, 0)
    a2, b2 = dd(b"a", [X, X], [160, 180]), dd(b"b", [X, Z], [160, 170])
    t4 = dd(b"t", [X, X, X, X], [5, 5, 5, 5])
    cases = [
        (st(b"a", [X], 160), st(b"b", [X, Z], 160), 100),
        (st(b"b", [X, Z], 160), st(b"a", [X], 160), 100),
        (a2, b2, 50), (b2, a2, 100),
        (dd(b"a", [X, X, X], [5, 100, 100]), t4, -1),
        (dd(b"a", [X, X, X], [100, 100, 5]), t4, 0),
        (dd(b"a", [X, X, X], [100, 5, 100]), t4, -1),
        (dd(b"a", [X], [5]), dd(b"t", [X, X], [160, 160]), 0),
        (dd(b"a", [X, X], [5, 5]), dd(b"t", [X, X, X], [160, 160, 160]), -1),
        (st(b"a", [bits((0, 1040))], 160), st(b"b", [bits((0, 340), (1100, 1800))], 160), 13),
        (st(b"a", [bits((0, 1040))], 160), st(b"b", [bits((0, 339), (1100, 1801))], 160), 12),
    ]
    for k, (x, y, want) in enumerate(cases):
        assert oracle.score(x, y) == want, k
        assert cmp_engine.compare_to(upload_oracle(cmp_engine, [x]), upload_oracle(cmp_engine, [y]))[0, 0] == want, k


def test_compare_autotuned(engine):
    # autotuning times the VALU kernel and the MFMA kernel at every chunk size
    # on the first call of a shape class (their scores must agree), returns
    # the scores, and reuses the fastest configuration afterwards
    import sdhash_amd
    img = synth.disk_image(35, 6 * MIB)
    frags = synth.fragments(36, img, 700)
    names = [b"img.%04dM" % k for k in range(6)]
    tgt = [oracle.digest_block(img[k * MIB:(k + 1) * MIB], 16384, n) for k, n in enumerate(names)]
    q = [oracle.digest_stream(f, b"frag%04d" % i) for i, f in enumerate(frags)]
    want = oracle.compare_to(q, tgt, threads=oracle.threads_available())
    gq = engine.new_set()
    assert np.all(gq.digest(frags, [d.name for d in q]) == 0)
    gt = engine.new_set()
    assert np.all(gt.digest([img[k * MIB:(k + 1) * MIB] for k in range(6)], names, dd_block=16384) == 0)
    engine.compare_method(sdhash_amd.CMP_AUTO)
    engine.autotune(True)
    try:
        first = engine.compare_to(gq, gt)
        table = engine.tuned()
        again = engine.compare_to(gq, gt)
        assert engine.tuned() == table
    finally:
        engine.autotune(False)
    assert np.array_equal(first, want) and np.array_equal(again, want)
    assert len(table) == 1 and table[0]["method"] in ("valu", "mfma") and table[0]["chunk"] in (1024, 2048, 4096,
                                                                                                   8192, 16384)
    # every query digest is the reference side against each of the 6 targets
    assert table[0]["instances"] == 6 * sum(d.bf_count for d in q) and table[0]["target_digests"] == 6

"""Pins the CPU oracle against checks that do not come from the oracle itself:
fingerprints of the reference's tables, FIPS-180 SHA-1 and RFC 4648 base64
vectors (through Python's hashlib/base64), closed forms (direct histogram
entropy, monotone/constant rank sequences), hand-traced known answers of the
popularity state machine (sdbf/sdbf_core.cc:51-89), hand-computed pair scores
and brute force.  CPU only."""
import base64
import hashlib
import os
import re

import numpy as np
import pytest

import oracle
from sdhash_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def fingerprint(values):
    h = 0
    for v in values:
        h = (h * 1000003 + int(v)) % (1 << 64)
    return "%016x" % h


def tables():
    import re, os
    path = os.path.join(os.path.dirname(oracle.__file__), "..", "sdhash_amd", "csrc", "sdbf_tables.h")
    src = open(path).read()

    def grab(name):
        body = src[src.index("#define " + name):]
        body = body[: body.index("\n\n")]
        return [int(x) for x in re.findall(r"\b\d+\b", body.split("\\", 1)[1])]
    return grab("SDBFX_ENTR64_RANKS_INIT"), grab("SDBFX_CUTOFFS256_INIT"), grab("SDBFX_ENTROPY64_INIT")


# fingerprints of the reference's constant tables (sdbf/sdbf_conf.cc:81-132,
# :145-154) and of the entropy table its build computes (sdbf/entr64.cc:14-21)
RANKS_FP = "6cfa7679e9b16c45"
CUTOFFS_FP = "fc0bf76f8a809444"
ENTROPY_FP = "b0e90c4ecc146f83"


def test_table_fingerprints():
    r, c, e = tables()
    assert len(r) == 1001 and len(c) == 149 and len(e) == 65
    assert fingerprint(r) == RANKS_FP
    assert fingerprint(c) == CUTOFFS_FP
    assert fingerprint(e) == ENTROPY_FP


def test_oracle_entropy_table_matches_reference_build():
    e = oracle.entropy_table()
    assert fingerprint(e) == ENTROPY_FP
    # closed form away from the two exactly-integral entries (c=1: 16000, c=8: 64000)
    for c in range(1, 65):
        p = c / 64
        assert abs(int(e[c]) - (-p * np.log2(p) / 6) * 1024000) < 1.0


# ---------------------------------------------------------------- SHA-1

FIPS_SHA1 = {
    b"abc": "a9993e364706816aba3e25717850c26c9cd0d89d",
    b"": "da39a3ee5e6b4b0d3255bfef95601890afd80709",
    b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq":
        "84983e441c3bd26ebaae4aa1f95129e5e54670f1",
}


def test_hashlib_is_fips_sha1():
    for msg, dig in FIPS_SHA1.items():
        assert hashlib.sha1(msg).hexdigest() == dig


def test_oracle_sha1_words_little_endian():
    rng = np.random.default_rng(1)
    for _ in range(50):
        blk = rng.integers(0, 256, 64, dtype=np.uint8)
        d = hashlib.sha1(blk.tobytes()).digest()
        want = np.frombuffer(d, dtype="<u4")
        assert np.array_equal(oracle.sha1_words(blk), want)


def test_bf_sha1_insert_set_semantics():
    rng = np.random.default_rng(2)
    bf = np.zeros(256, np.uint8)
    seen = set()
    for _ in range(400):
        h = rng.integers(0, 1 << 32, 5, dtype=np.uint64).astype(np.uint32)
        want = 0
        for w in h:
            pos = int(w) & 0x7FF
            if pos not in seen:
                want += 1
                seen.add(pos)
        assert oracle.bf_sha1_insert(bf, h) == want
    bits = np.unpackbits(bf, bitorder="little")
    assert set(np.flatnonzero(bits)) == seen


# ---------------------------------------------------------------- ranks

def direct_ranks(buf: np.ndarray) -> np.ndarray:
    """Window entropy as a direct 64-byte histogram sum per window."""
    r_tab, _, e_tab = tables()
    r_tab = np.array(r_tab)
    e_tab = np.array(e_tab, dtype=np.int64)
    n = buf.size
    out = np.zeros(n, np.uint16)
    if n <= 64:
        return out
    for o in range(n - 64):
        cnt = np.bincount(buf[o:o + 64], minlength=256)
        out[o] = r_tab[int(e_tab[cnt].sum()) >> 10]
    return out


@pytest.mark.parametrize("seed,n", [(1, 600), (2, 5000), (3, 9000)])
def test_ranks_equal_direct_histogram(seed, n):
    buf = synth.segment_mix(np.random.Generator(np.random.PCG64(seed)), n)
    assert np.array_equal(oracle.ranks(buf), direct_ranks(buf))


def test_ranks_uniform_and_zero_windows():
    # 64 distinct bytes: entropy 64*16000 = 1024000 -> bin 1000 -> rank 0
    buf = np.tile(np.arange(64, dtype=np.uint8), 3)
    assert np.all(oracle.ranks(buf) == 0)
    # all zeros: entropy E[64] = 0 -> rank 0
    assert np.all(oracle.ranks(np.zeros(300, np.uint8)) == 0)


# ---------------------------------------------------------------- scores

N = 200          # iterations 0..135; ranks at 136..199 are zero (end of chunk)
NIT = N - 64


def ranks_of(head):
    r = np.zeros(N, np.uint16)
    r[:NIT] = head
    return r


def test_scores_increasing_ranks_credit_window_start():
    s = oracle.scores(ranks_of(100 + np.arange(NIT)))
    want = np.zeros(N, np.uint16)
    want[:NIT] = 1
    assert np.array_equal(s, want)


def test_scores_decreasing_ranks_credit_window_end():
    # window [i, i+63]: minimum at min(i+63, 135); from i=73 on, every full
    # rescan lands on 135, the last non-zero rank
    s = oracle.scores(ranks_of(1000 - np.arange(NIT)))
    want = np.zeros(N, np.uint16)
    want[63:135] = 1          # i = 0..71
    want[135] = 1 + 63        # i = 72, then i = 73..135
    assert np.array_equal(s, want)


def test_scores_constant_region_hand_trace():
    # i=0: rescan, equal chain -> 63.  i=1..71: cheap slide, r[i+64]==r[m] ->
    # m=i+64 (position 64 is never credited).  i=72..135: r[i+64]=0 breaks the
    # slide; every rescan ends on 135.
    s = oracle.scores(ranks_of(np.full(NIT, 500)))
    want = np.zeros(N, np.uint16)
    want[63] = 1
    want[65:135] = 1
    want[135] = 1 + 64
    assert np.array_equal(s, want)


def test_scores_cheap_slide_misses_smaller_rank():
    # a single smaller rank at position 64 is inside windows 1..64, but the
    # cheap slide only looks one past the window, so it is never credited:
    # the result equals the constant-region trace above
    head = np.full(NIT, 500)
    head[64] = 100
    s = oracle.scores(ranks_of(head))
    want = np.zeros(N, np.uint16)
    want[63] = 1
    want[65:135] = 1
    want[135] = 65
    assert np.array_equal(s, want)


def test_scores_zero_leading_windows_never_score():
    head = 100 + np.arange(NIT)
    head[:10] = 0
    s = oracle.scores(ranks_of(head))
    want = np.zeros(N, np.uint16)
    want[10:NIT] = 1
    assert np.array_equal(s, want)


def test_scores_sum_is_steps_with_nonzero_election():
    buf = synth.segment_mix(np.random.Generator(np.random.PCG64(9)), 20000)
    r = oracle.ranks(buf)
    s, h = oracle.scores(r, with_histo=True)
    assert s.sum() <= buf.size - 64
    assert h.sum() == buf.size - 64
    assert np.array_equal(h, np.bincount(s[: buf.size - 64], minlength=66)[:66])


# ---------------------------------------------------------------- block threshold

@pytest.mark.parametrize("hist,k,allowed", [
    ({}, 15, 192),                         # loop never breaks
    ({65: 200}, 65, 192),
    ({20: 100, 17: 100}, 17, 92),
    ({30: 192, 16: 5}, 16, 0),             # sum reaches exactly 192 at k=30
    ({40: 50, 25: 150}, 25, 142),
])
def test_block_threshold_hand_cases(hist, k, allowed):
    h = np.zeros(66, np.int32)
    for kk, v in hist.items():
        h[kk] = v
    assert oracle.block_threshold(h) == (k, allowed)


# ---------------------------------------------------------------- compare

def test_bitcount_cut_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(200):
        a = rng.integers(0, 256, 256, dtype=np.uint8)
        b = rng.integers(0, 256, 256, dtype=np.uint8) & rng.integers(0, 256, 256, dtype=np.uint8)
        full = int(np.unpackbits(a & b).sum())
        assert oracle.bitcount_cut(a, b, 0, 0) == full
        cut = int(rng.integers(2, 1250))
        p = [int(np.unpackbits((a & b)[:k]).sum()) for k in (32, 64, 128)]
        ok = 8 * p[0] + 48 >= cut and 4 * p[1] + 48 >= cut and 2 * p[2] + 48 >= cut
        assert oracle.bitcount_cut(a, b, cut, 48) == (full if ok else 0)


def bits(*ranges):
    v = np.zeros(2048, np.uint8)
    for lo, hi in ranges:
        v[lo:hi] = 1
    return np.packbits(v, bitorder="little")


def one_filter(name, f, ec=160):
    return oracle.Digest(name, f, 1, ec, None, 1000, 0)


def test_pair_score_hand_computed():
    # s1=s2=160 -> cut = CUTOFFS256[4096//320 = 12] = 240;
    # match 500, max_est 600 -> (500-240)/(600-240) = 0.7222 -> 72
    a = one_filter(b"a", bits((0, 600)))
    b = one_filter(b"b", bits((0, 500), (1500, 1600)))
    assert oracle.score(a, b) == 72
    assert oracle.score(b, a) == 72


def test_pair_score_early_exit_zeroes_a_full_match():
    # 10 common bits in the first 32 bytes: 8*10+48 = 128 < 240, so the
    # (otherwise identical) filters score 0 instead of 100
    f = bits((0, 10), (1024, 1624))
    assert oracle.score(one_filter(b"a", f), one_filter(b"b", f.copy())) == 0
    # 30 bits pass the 32-byte check (8*30+48 = 288) but fail the 64-byte one
    # (4*30+48 = 168 < 240)
    f = bits((0, 30), (1024, 1624))
    assert oracle.score(one_filter(b"a", f), one_filter(b"b", f.copy())) == 0
    # 90 bits in the first 128 bytes fail the last check (2*90+48 = 228 < 240)
    f = bits((0, 90), (1024, 1624))
    assert oracle.score(one_filter(b"a", f), one_filter(b"b", f.copy())) == 0
    # 100 bits pass all three (248 >= 240): identical filters score 100
    g = bits((0, 100), (1024, 1624))
    assert oracle.score(one_filter(b"a", g), one_filter(b"b", g.copy())) == 100


def test_pair_score_sparse_target_is_minus_one():
    a = one_filter(b"a", bits((0, 600)), ec=100)
    b = oracle.Digest(b"b", np.concatenate([bits((0, 600)), bits((0, 600))]), 2, 0,
                      np.array([5, 5], np.uint16), 1000, 16384)
    assert oracle.score(a, b) == -1


def test_self_compare_is_100_and_roundtrip():
    f = synth.file_set(5, 2, 1 << 18)
    d = oracle.digest_stream(f[0], b"f0")
    assert oracle.score(d, d) == 100
    blk = oracle.digest_block(f[1], 16384, b"f1")
    assert oracle.score(blk, blk) == 100


# ---------------------------------------------------------------- text form

def parse_stream(text):
    parts = text.rstrip("\n").split(":")
    assert parts[0] == "sdbf" and parts[1] == "03"
    name_len = int(parts[2])
    name = parts[3]
    assert len(name) == name_len
    return dict(name=name, size=int(parts[4]), bf_size=int(parts[6]), k=int(parts[7]), mask=parts[8],
                max_elem=int(parts[9]), bf_count=int(parts[10]), last=int(parts[11]),
                filters=base64.b64decode(parts[12]))


def test_stream_to_string_grammar_and_base64():
    f = synth.file_set(6, 1, 300000)[0]
    d = oracle.digest_stream(f, b"file.bin")
    t = parse_stream(d.to_string())
    assert t["name"] == "file.bin" and t["size"] == f.size and t["mask"] == "7ff"
    assert (t["bf_size"], t["k"], t["max_elem"]) == (256, 5, 160)
    assert t["bf_count"] == d.bf_count and t["last"] == d.last_count
    assert t["filters"] == d.filters.tobytes()


def test_block_to_string_grammar_and_base64():
    f = synth.file_set(7, 1, 100000)[0]
    d = oracle.digest_block(f, 16384, b"img")
    parts = d.to_string().rstrip("\n").split(":")
    assert parts[:12] == ["sdbf-dd", "03", "3", "img", str(f.size), "sha1", "256", "5", "7ff", "192",
                          str(d.bf_count), "16384"]
    rest = parts[12:]
    assert len(rest) == 2 * d.bf_count
    for i in range(d.bf_count):
        assert rest[2 * i] == "%02x" % d.elem_counts[i]
        assert len(rest[2 * i + 1]) == 344
        assert base64.b64decode(rest[2 * i + 1]) == d.filters[256 * i:256 * (i + 1)].tobytes()


# ---------------------------------------------------------------- digest invariants

def test_stream_digest_invariants():
    f = synth.file_set(8, 1, 1 << 20)[0]
    d = oracle.digest_stream(f)
    assert d.bf_count >= 1 and 20 <= d.last_count <= 160 or d.bf_count == 1
    # every selected feature's bits are in its filter: the filters' OR covers
    # the bits of all positions with score > 16 in order (checked by hashing)
    r = oracle.ranks(f)
    s = oracle.scores(r)
    allbits = np.zeros(2048, bool)
    for i in np.flatnonzero(s[: f.size - 64] > 16)[:50]:
        h = oracle.sha1_words(f[i:i + 64])
        for w in h:
            allbits[int(w) & 0x7FF] = True
    first = np.unpackbits(d.filters[:256], bitorder="little").astype(bool)
    assert np.all(first[allbits])


def test_block_digest_elem_counts_bounded():
    f = synth.disk_image(11, 16384 * 8 + 700)
    d = oracle.digest_block(f, 16384)
    assert d.bf_count == 9
    assert np.all(d.elem_counts <= 192)
    assert oracle.digest_block(f[:16384 * 8 + 300], 16384).bf_count == 8   # tail < 512 dropped


def test_cutoff_table_monotone_in_size():
    # The MFMA comparison rules a pair out when its full AND-popcount is at
    # most cut(s1min + s2min); that is a lower bound of every cut(s1 + s2) in
    # the tile only if cut(s) = CUTOFFS256[4096 // s] never decreases for
    # s >= 32 (both filters have >= 16 elements) and sums above 1024 map to
    # 1250, the table's largest entry.
    import re
    text = open(os.path.join(ROOT, "sdhash_amd", "csrc", "sdbf_tables.h")).read()
    body = text.split("#define SDBFX_CUTOFFS256_INIT", 1)[1].split("\n\n", 1)[0]
    table = [int(x) for x in re.findall(r"\d+", body)]
    cut = [table[4096 // s] for s in range(28, 1025)]
    assert all(a <= b for a, b in zip(cut, cut[1:]))
    assert max(table) == 1250 and cut[-1] <= 1250


# ---- fast mode (sdbf::fast, sdbf/sdbf_class.cc:426-439)

CUTOFFS64_FP = "06a3db5f5e2e7675"   # sdbf/sdbf_conf.cc:156-165
KIB = 1024


def test_cutoffs64_fingerprint():
    path = os.path.join(ROOT, "sdhash_amd", "csrc", "sdbf_tables.h")
    src = open(path).read()
    body = src[src.index("#define SDBFX_CUTOFFS64_INIT"):]
    body = body[: body.index("\n\n")]
    c = [int(x) for x in re.findall(r"\b\d+\b", body.split("\\", 1)[1])]
    assert len(c) == 149 and fingerprint(c) == CUTOFFS64_FP


def test_fold_is_or_of_quarters():
    rng = np.random.Generator(np.random.PCG64(71))
    f = (rng.random(5 * 256) < 0.3).astype(np.uint8) * rng.integers(0, 256, 5 * 256, dtype=np.uint8)
    d = oracle.Digest(b"x", f, 5, 100, None, 5000, 0)
    g = d.folded()
    a = f.reshape(5, 4, 64)
    want = a[:, 0] | a[:, 1] | a[:, 2] | a[:, 3]
    got = g.filters.reshape(5, 256)
    assert np.array_equal(got[:, :64], want)
    assert np.array_equal(got[:, 64:], f.reshape(5, 256)[:, 64:])
    assert np.array_equal(g.hamming[:5], np.unpackbits(want, axis=1).sum(axis=1))
    assert np.array_equal(d.filters, f)   # the original is left as it was


def test_bitcount_cut_64_brute_force():
    rng = np.random.Generator(np.random.PCG64(72))
    for _ in range(200):
        a = rng.integers(0, 256, 64, dtype=np.uint8) & rng.integers(0, 256, 64, dtype=np.uint8)
        b = rng.integers(0, 256, 64, dtype=np.uint8)
        full = int(np.unpackbits(a & b).sum())
        half = int(np.unpackbits(a[:32] & b[:32]).sum())
        assert oracle.bitcount_cut_64(a, b, 0, 0) == full
        cut = int(rng.integers(0, 200))
        want = 0 if cut > 0 and 2 * half + 48 < cut else full
        assert oracle.bitcount_cut_64(a, b, cut, 48) == want


def test_fast_scores_closed_forms():
    # a digest against an identical copy: every folded filter matches itself
    # fully (fraction 1) -> 100; against bit-disjoint filters -> 0
    files = synth.file_set(73, 2, 200 * KIB)
    d = oracle.digest_stream(files[0], b"a").folded()
    e = oracle.Digest(b"b", d.filters.copy(), d.bf_count, d.last_count, None, d.orig_size, 0,
                      d.hamming.copy(), True)
    assert oracle.compare_all([d, e])[0] == 100
    # disjoint: the same filters with their folded halves moved to bits that
    # the other side never sets
    lo = d.filters.reshape(-1, 256).copy()
    hi = lo.copy()
    lo[:, :64] &= 0x0F
    hi[:, :64] &= 0xF0
    x = oracle.Digest(b"x", lo.reshape(-1), d.bf_count, d.last_count, None, 1, 0,
                      np.unpackbits(lo[:, :64], axis=1).sum(axis=1).astype(np.uint16), True)
    y = oracle.Digest(b"y", hi.reshape(-1), d.bf_count, d.last_count, None, 1, 0,
                      np.unpackbits(hi[:, :64], axis=1).sum(axis=1).astype(np.uint16), True)
    assert oracle.compare_all([x, y])[0] == 0

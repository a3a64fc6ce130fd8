"""ctypes front end of the CPU oracle (oracle/sdbf_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's cpu_baseline leg.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# tests/test_oracle_mutants.py points this at builds with planted mistakes
_SO = os.environ.get("SDBF_ORACLE_SO") or os.path.join(_HERE, "liboracle.so")
_lib = None


REF_LZ4 = os.path.join(_HERE, "_ref", "lz4check")


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    # the reference's own LZ4 (index files), where the sdhash tree is present
    if os.path.exists("/root/reference/lz4/lz4.cc"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        P = C.c_void_p
        L.orc_entropy_table.argtypes = [P]
        L.orc_gen_ranks.argtypes = [P, C.c_uint64, P]
        L.orc_gen_scores.argtypes = [P, C.c_uint64, P, P]
        L.orc_bf_sha1_insert.argtypes = [P, P]
        L.orc_bf_sha1_insert.restype = C.c_uint32
        L.orc_sha1_words.argtypes = [P, P]
        L.orc_digest_stream.argtypes = [P, C.c_uint64, C.POINTER(P), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint32)]
        L.orc_block_count.argtypes = [C.c_uint64, C.c_uint32]
        L.orc_block_count.restype = C.c_uint64
        L.orc_digest_block.argtypes = [P, C.c_uint64, C.c_uint32, P, P, C.c_int]
        L.orc_block_threshold.argtypes = [P, C.POINTER(C.c_uint32), C.POINTER(C.c_int32)]
        L.orc_hamming.argtypes = [P, C.c_uint32, P]
        L.orc_bitcount_cut_256.argtypes = [P, P, C.c_uint32, C.c_int32]
        L.orc_bitcount_cut_256.restype = C.c_uint32
        L.orc_bitcount_cut_64.argtypes = [P, P, C.c_uint32, C.c_int32]
        L.orc_bitcount_cut_64.restype = C.c_uint32
        L.orc_fold_fast.argtypes = [P, C.c_uint32, P]
        L.orc_max_score.argtypes = [P, C.c_uint32, P]
        L.orc_max_score.restype = C.c_double
        L.orc_score.argtypes = [P, P]
        L.orc_score.restype = C.c_int
        L.orc_compare_all.argtypes = [P, C.c_uint32, P, C.c_int]
        L.orc_compare_to.argtypes = [P, C.c_uint32, P, C.c_uint32, P, C.c_int]
        L.orc_compare_to_sampled.argtypes = [P, C.c_uint32, P, C.c_uint32, C.c_uint32, P, C.c_int]
        L.orc_score_sampled.argtypes = [P, P, C.c_uint32]
        L.orc_score_sampled.restype = C.c_int
        L.orc_digest_stream_many.argtypes = [C.c_uint32, P, P, P, P, P, P, C.c_int]
        L.orc_free.argtypes = [P]
        L.orc_b64encode.argtypes = [P, C.c_uint64, P]
        L.orc_b64encode.restype = C.c_uint64
        L.orc_to_string.argtypes = [P, C.c_uint64, C.c_uint32, P]
        L.orc_to_string.restype = C.c_uint64
        L.orc_elem_count.argtypes = [P, C.c_uint64]
        L.orc_elem_count.restype = C.c_int32
        L.orc_threads_available.restype = C.c_int
        L.orc_stream_hash_scored.argtypes = [P, P, C.c_uint64, C.c_uint64, C.POINTER(P), C.POINTER(C.c_uint32),
                                             C.POINTER(C.c_uint32)]
        L.orc_block_digest_scored.argtypes = [P, P, C.c_uint64, C.c_uint32, P, P]
        L.orc_block_hash_scored.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_int32, P]
        L.orc_block_hash_scored.restype = C.c_uint32
        U64P = C.POINTER(C.c_uint64)
        L.orc_index_mask.argtypes = [C.c_uint64]
        L.orc_index_mask.restype = C.c_uint32
        L.orc_index_insert.argtypes = [P, C.c_uint64, U64P, P]
        L.orc_index_query.argtypes = [P, C.c_uint64, P]
        L.orc_digest_stream_indexed.argtypes = [P, C.c_uint64, P, C.c_uint64, U64P, C.POINTER(P),
                                                C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_stream_hash_scored_indexed.argtypes = [P, P, C.c_uint64, C.c_uint64, P, C.c_uint64, U64P,
                                                     C.POINTER(P), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_digest_block_indexed.argtypes = [P, C.c_uint64, C.c_uint32, P, C.c_uint64, U64P, P, P, C.c_uint32,
                                               P, P, P]
        L.orc_block_hash_scored_indexed.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_int32, P, P, C.c_uint64,
                                                    U64P, P, P, C.c_uint32, P]
        L.orc_block_hash_scored_indexed.restype = C.c_uint32
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _u8(buf) -> np.ndarray:
    if isinstance(buf, np.ndarray):
        return np.ascontiguousarray(buf, dtype=np.uint8)
    return np.frombuffer(bytes(buf), dtype=np.uint8)


def entropy_table() -> np.ndarray:
    out = np.zeros(65, dtype=np.uint64)
    lib().orc_entropy_table(_ptr(out))
    return out


def ranks(buf) -> np.ndarray:
    b = _u8(buf)
    r = np.zeros(b.size, dtype=np.uint16)
    lib().orc_gen_ranks(_ptr(b), b.size, _ptr(r))
    return r


def scores(rk: np.ndarray, with_histo: bool = False):
    rk = np.ascontiguousarray(rk, dtype=np.uint16)
    s = np.zeros(rk.size, dtype=np.uint16)
    h = np.zeros(66, dtype=np.int32)
    lib().orc_gen_scores(_ptr(rk), rk.size, _ptr(s), _ptr(h) if with_histo else None)
    return (s, h) if with_histo else s


def sha1_words(block64) -> np.ndarray:
    b = _u8(block64)
    assert b.size >= 64
    h = np.zeros(5, dtype=np.uint32)
    lib().orc_sha1_words(_ptr(b), _ptr(h))
    return h


def bf_sha1_insert(bf: np.ndarray, h: np.ndarray) -> int:
    return int(lib().orc_bf_sha1_insert(_ptr(bf), _ptr(np.ascontiguousarray(h, np.uint32))))


def block_threshold(histo) -> tuple[int, int]:
    h = np.ascontiguousarray(histo, dtype=np.int32)
    k = C.c_uint32()
    a = C.c_int32()
    lib().orc_block_threshold(_ptr(h), C.byref(k), C.byref(a))
    return k.value, a.value


def bitcount_cut(a, b, cut_off: int, slack: int) -> int:
    return int(lib().orc_bitcount_cut_256(_ptr(_u8(a)), _ptr(_u8(b)), cut_off, slack))


def bitcount_cut_64(a, b, cut_off: int, slack: int) -> int:
    return int(lib().orc_bitcount_cut_64(_ptr(_u8(a)), _ptr(_u8(b)), cut_off, slack))


class _CDigest(C.Structure):
    _fields_ = [("bf", C.c_void_p), ("elem_counts", C.c_void_p), ("hamming", C.c_void_p),
                ("name", C.c_char_p), ("bf_count", C.c_uint32), ("last_count", C.c_uint32),
                ("max_elem", C.c_uint32), ("fast", C.c_uint32)]


@dataclass
class Digest:
    """One sdbf digest as the oracle computes it."""
    name: bytes
    filters: np.ndarray            # uint8 [bf_count * 256]
    bf_count: int
    last_count: int = 0            # stream mode
    elem_counts: np.ndarray | None = None   # dd mode
    orig_size: int = 0
    dd_block_size: int = 0
    hamming: np.ndarray = field(default=None)
    fast: bool = False             # filters folded to 64 bytes (sdbf::fast)
    _c: _CDigest | None = field(default=None, repr=False)

    @property
    def max_elem(self) -> int:
        return 192 if self.elem_counts is not None else 160

    def cstruct(self) -> _CDigest:
        if self._c is None:
            self.filters = np.ascontiguousarray(self.filters, dtype=np.uint8)
            if not self.fast:
                self.hamming = np.zeros(max(self.bf_count, 1), dtype=np.uint16)
                lib().orc_hamming(_ptr(self.filters), self.bf_count, _ptr(self.hamming))
            ec = None
            if self.elem_counts is not None:
                self.elem_counts = np.ascontiguousarray(self.elem_counts, dtype=np.uint16)
                ec = self.elem_counts.ctypes.data
            self._c = _CDigest(self.filters.ctypes.data, ec, self.hamming.ctypes.data,
                               self.name, self.bf_count, self.last_count, self.max_elem, int(self.fast))
        return self._c

    def folded(self) -> "Digest":
        """A copy in fast mode (sdbf::fast): every filter folded to 64 bytes
        in place, Hamming weights of the folded bytes."""
        f = np.array(self.filters, dtype=np.uint8, copy=True)
        ham = np.zeros(max(self.bf_count, 1), dtype=np.uint16)
        lib().orc_fold_fast(_ptr(f), self.bf_count, _ptr(ham))
        ec = None if self.elem_counts is None else np.array(self.elem_counts, dtype=np.uint16, copy=True)
        return Digest(self.name, f, self.bf_count, self.last_count, ec, self.orig_size, self.dd_block_size,
                      ham, True)

    def elem_count(self, i: int) -> int:
        return int(lib().orc_elem_count(C.byref(self.cstruct()), i))

    def to_string(self) -> str:
        c = self.cstruct()
        n = lib().orc_to_string(C.byref(c), self.orig_size, self.dd_block_size, None)
        buf = C.create_string_buffer(int(n) + 1)
        m = lib().orc_to_string(C.byref(c), self.orig_size, self.dd_block_size, buf)
        return buf.raw[:m].decode("ascii")


def digest_stream(buf, name: bytes = b"x") -> Digest:
    b = _u8(buf)
    p = C.c_void_p()
    n = C.c_uint32()
    last = C.c_uint32()
    rc = lib().orc_digest_stream(_ptr(b), b.size, C.byref(p), C.byref(n), C.byref(last))
    if rc != 0:
        raise ValueError(rc)
    f = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n.value * 256,)).copy()
    lib().orc_free(p)
    return Digest(name, f, n.value, last.value, None, b.size, 0)


def digest_block(buf, dd: int = 16384, name: bytes = b"x", threads: int = 1) -> Digest:
    b = _u8(buf)
    if b.size < 512:
        raise ValueError(-3)
    cnt = int(lib().orc_block_count(b.size, dd))
    f = np.zeros(cnt * 256, dtype=np.uint8)
    ec = np.zeros(max(cnt, 1), dtype=np.uint16)
    rc = lib().orc_digest_block(_ptr(b), b.size, dd, _ptr(f), _ptr(ec), threads)
    if rc != 0:
        raise ValueError(rc)
    return Digest(name, f, cnt, 0, ec[:cnt], b.size, dd)


def digest_stream_many(bufs, names=None, threads: int = 1) -> list[Digest]:
    arrs = [_u8(b) for b in bufs]
    n = len(arrs)
    ptrs = (C.c_void_p * n)(*[a.ctypes.data for a in arrs])
    lens = np.array([a.size for a in arrs], dtype=np.uint64)
    fptr = (C.c_void_p * n)()
    cnt = np.zeros(n, dtype=np.uint32)
    last = np.zeros(n, dtype=np.uint32)
    st = np.zeros(n, dtype=np.int32)
    lib().orc_digest_stream_many(n, ptrs, _ptr(lens), fptr, _ptr(cnt), _ptr(last), _ptr(st),
                                 threads)
    out = []
    for i in range(n):
        if st[i] != 0:
            out.append(None)
            continue
        f = np.ctypeslib.as_array(C.cast(fptr[i], C.POINTER(C.c_uint8)),
                                  shape=(int(cnt[i]) * 256,)).copy()
        lib().orc_free(fptr[i])
        nm = names[i] if names is not None else b"f%05d" % i
        out.append(Digest(nm, f, int(cnt[i]), int(last[i]), None, arrs[i].size, 0))
    return out


def _scored_inputs(scores, words):
    sc = np.ascontiguousarray(scores, dtype=np.uint16)
    w = np.ascontiguousarray(words, dtype=np.uint32).reshape(-1)
    assert w.size == 5 * sc.size
    return sc, w


def stream_hash_scored(scores, words, unit: int = 32 << 20):
    """The stream path (unit loop, gen_chunk_hash, trim) over given
    per-position scores and SHA-1 words (words: [n, 5]); returns
    (filters uint8[bf_count*256], bf_count, last_count)."""
    sc, w = _scored_inputs(scores, words)
    p = C.c_void_p()
    n = C.c_uint32()
    last = C.c_uint32()
    lib().orc_stream_hash_scored(_ptr(sc), _ptr(w), sc.size, unit, C.byref(p), C.byref(n), C.byref(last))
    f = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n.value * 256,)).copy()
    lib().orc_free(p)
    return f, n.value, last.value


def block_digest_scored(scores, words, dd: int):
    """The block path (per-block histogram, threshold, gen_block_hash, tail
    rule) over given per-position scores and SHA-1 words; returns
    (filters uint8[count*256], elem_counts uint16[count])."""
    sc, w = _scored_inputs(scores, words)
    cnt = int(lib().orc_block_count(sc.size, dd))
    f = np.zeros(max(cnt, 1) * 256, dtype=np.uint8)
    ec = np.zeros(max(cnt, 1), dtype=np.uint16)
    lib().orc_block_digest_scored(_ptr(sc), _ptr(w), sc.size, dd, _ptr(f), _ptr(ec))
    return f[: cnt * 256], ec[:cnt]


def block_hash_scored(scores, words, threshold: int, allowed: int):
    """One gen_block_hash call over positions [0, len(scores) - 64) with the
    given threshold and allowance; returns (filter uint8[256], hash_cnt)."""
    sc, w = _scored_inputs(scores, words)
    bf = np.zeros(256, dtype=np.uint8)
    n = lib().orc_block_hash_scored(_ptr(sc), _ptr(w), sc.size, threshold, allowed, _ptr(bf))
    return bf, int(n)


def max_score(ref: Digest, i: int, tgt: Digest) -> float:
    return float(lib().orc_max_score(C.byref(ref.cstruct()), i, C.byref(tgt.cstruct())))


def score(a: Digest, b: Digest, sample: int = 0) -> int:
    """a.compare(b, sample) (sdbf_score); INT32_MIN where the reference aborts
    on an infinite score."""
    return int(lib().orc_score_sampled(C.byref(a.cstruct()), C.byref(b.cstruct()), sample))


def _array(ds):
    arr = (_CDigest * len(ds))()
    for i, d in enumerate(ds):
        arr[i] = d.cstruct()
    return arr


def compare_all(ds: list[Digest], threads: int = 1) -> np.ndarray:
    """Scores of all pairs i<j in row-major order."""
    n = len(ds)
    out = np.zeros(n * (n - 1) // 2, dtype=np.int32)
    if n > 1:
        lib().orc_compare_all(_array(ds), n, _ptr(out), threads)
    return out


def compare_to(q: list[Digest], t: list[Digest], threads: int = 1, sample: int = 0) -> np.ndarray:
    out = np.zeros((len(q), len(t)), dtype=np.int32)
    if len(q) and len(t):
        lib().orc_compare_to_sampled(_array(q), len(q), _array(t), len(t), sample, _ptr(out), threads)
    return out


def threads_available() -> int:
    return int(lib().orc_threads_available())


# ------------------------------------------------------------------ index


class Index:
    """An index as the oracle holds it: bloom_filter(size, 5, 0, 0.01)
    (sdbf/bloom_filter.cc:52-71) -- `size` bytes and an element count."""

    def __init__(self, size: int = 64 << 20, bits=None, elem_count: int = 0):
        self.size = int(size)
        self.bf = np.zeros(self.size, np.uint8) if bits is None else np.array(bits, np.uint8, copy=True)
        assert self.bf.size == self.size
        self.elem_count = int(elem_count)

    @property
    def mask(self) -> int:
        return int(lib().orc_index_mask(self.size))

    def copy(self) -> "Index":
        return Index(self.size, self.bf, self.elem_count)

    def insert(self, h) -> bool:
        w = np.ascontiguousarray(h, np.uint32)
        e = C.c_uint64(self.elem_count)
        r = lib().orc_index_insert(_ptr(self.bf), self.size, C.byref(e), _ptr(w))
        self.elem_count = e.value
        return bool(r)

    def query(self, h) -> bool:
        w = np.ascontiguousarray(h, np.uint32)
        return bool(lib().orc_index_query(_ptr(self.bf), self.size, _ptr(w)))


def _search_arrays(search):
    n = len(search)
    bfs = (C.c_void_p * max(n, 1))(*[x.bf.ctypes.data for x in search])
    sizes = np.array([x.size for x in search] or [0], np.uint64)
    return n, bfs, sizes


def digest_stream_indexed(buf, index: Index, name: bytes = b"x") -> Digest:
    """Stream digest with info->index = index (updated in place)."""
    b = _u8(buf)
    p = C.c_void_p()
    n = C.c_uint32()
    last = C.c_uint32()
    e = C.c_uint64(index.elem_count)
    rc = lib().orc_digest_stream_indexed(_ptr(b), b.size, _ptr(index.bf), index.size, C.byref(e), C.byref(p),
                                         C.byref(n), C.byref(last))
    if rc != 0:
        raise ValueError(rc)
    index.elem_count = e.value
    f = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n.value * 256,)).copy()
    lib().orc_free(p)
    return Digest(name, f, n.value, last.value, None, b.size, 0)


def digest_block_indexed(buf, dd: int = 16384, name: bytes = b"x", index: Index | None = None, search=()):
    """Block digest with an index_info: index (info->index, updated) and/or
    the searched indexes; returns (Digest, matches uint32[blocks, len(search)])."""
    b = _u8(buf)
    if b.size < 512:
        raise ValueError(-3)
    cnt = int(lib().orc_block_count(b.size, dd))
    f = np.zeros(cnt * 256, dtype=np.uint8)
    ec = np.zeros(max(cnt, 1), dtype=np.uint16)
    ns, bfs, sizes = _search_arrays(list(search))
    m = np.zeros(max(cnt * ns, 1), np.uint32)
    e = C.c_uint64(index.elem_count if index else 0)
    rc = lib().orc_digest_block_indexed(_ptr(b), b.size, dd, _ptr(index.bf) if index else None,
                                        index.size if index else 0, C.byref(e), bfs, _ptr(sizes), ns, _ptr(f),
                                        _ptr(ec), _ptr(m))
    if rc != 0:
        raise ValueError(rc)
    if index:
        index.elem_count = e.value
    return Digest(name, f, cnt, 0, ec[:cnt], b.size, dd), m[:cnt * ns].reshape(cnt, ns)


def stream_hash_scored_indexed(scores, words, index: Index, unit: int = 32 << 20):
    """stream_hash_scored with the index gate; index updated in place."""
    sc, w = _scored_inputs(scores, words)
    p = C.c_void_p()
    n = C.c_uint32()
    last = C.c_uint32()
    e = C.c_uint64(index.elem_count)
    lib().orc_stream_hash_scored_indexed(_ptr(sc), _ptr(w), sc.size, unit, _ptr(index.bf), index.size, C.byref(e),
                                         C.byref(p), C.byref(n), C.byref(last))
    index.elem_count = e.value
    f = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n.value * 256,)).copy()
    lib().orc_free(p)
    return f, n.value, last.value


def block_hash_scored_indexed(scores, words, threshold: int, allowed: int, index: Index | None = None, search=()):
    """block_hash_scored with an index_info; returns (filter, hash_cnt,
    matches uint32[len(search)])."""
    sc, w = _scored_inputs(scores, words)
    bf = np.zeros(256, dtype=np.uint8)
    ns, bfs, sizes = _search_arrays(list(search))
    m = np.zeros(max(ns, 1), np.uint32)
    e = C.c_uint64(index.elem_count if index else 0)
    r = lib().orc_block_hash_scored_indexed(_ptr(sc), _ptr(w), sc.size, threshold, allowed, _ptr(bf),
                                            _ptr(index.bf) if index else None, index.size if index else 0,
                                            C.byref(e), bfs, _ptr(sizes), ns, _ptr(m))
    if index:
        index.elem_count = e.value
    return bf, int(r), m[:ns]

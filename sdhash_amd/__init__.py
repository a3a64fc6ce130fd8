"""MI355X sdbf engine: Python front end over the C ABI of include/sdbfx.h.

The compute path is libsdbfx.so (gfx950 kernels); there is no CPU fallback:
if the library is missing or no GPU is present, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsdbfx.so")

# C ABI entry points (all declared in include/sdbfx.h)
SYMBOLS = (
    "sdbfx_ctx_create", "sdbfx_ctx_destroy", "sdbfx_last_error", "sdbfx_sync", "sdbfx_ctx_scratch_bytes",
    "sdbfx_ctx_trim",
    "sdbfx_set_create", "sdbfx_set_destroy", "sdbfx_set_size", "sdbfx_set_filter_count",
    "sdbfx_set_digest", "sdbfx_set_digest_device", "sdbfx_set_add", "sdbfx_set_info",
    "sdbfx_set_read", "sdbfx_set_to_text", "sdbfx_set_add_text", "sdbfx_set_fold", "sdbfx_set_folded",
    "sdbfx_set_fold_digest", "sdbfx_set_digest_folded",
    "sdbfx_compare_all", "sdbfx_compare_to", "sdbfx_compare_to_sampled", "sdbfx_compare_to_range",
    "sdbfx_device_alloc", "sdbfx_device_free", "sdbfx_upload", "sdbfx_set_profiling", "sdbfx_kernel_time",
    "sdbfx_reset_kernel_times", "sdbfx_set_compare_method", "sdbfx_digest_blocks_noalloc",
    "sdbfx_serve_create", "sdbfx_serve_destroy", "sdbfx_serve_error", "sdbfx_serve_ref_size",
    "sdbfx_serve_submit", "sdbfx_serve_poll", "sdbfx_serve_wait", "sdbfx_serve_stats", "sdbfx_serve_admission",
    "sdbfx_index_create", "sdbfx_index_destroy", "sdbfx_index_size", "sdbfx_index_read", "sdbfx_index_write",
    "sdbfx_set_digest_indexed", "sdbfx_set_digest_device_indexed", "sdbfx_set_autotune", "sdbfx_autotune_table",
    "sdbfx_graph_blocks_create", "sdbfx_graph_launch", "sdbfx_graph_destroy",
)

# comparison kernels (sdbfx_set_compare_method)
CMP_AUTO, CMP_VALU, CMP_MFMA, CMP_MFMA_I8 = 0, 1, 2, 3

_lib = None


class SdbfxError(RuntimeError):
    def __init__(self, msg, status=-1):
        super().__init__(msg)
        self.status = status


EBUDGET = -4   # a call's scratch would exceed the context's memory budget


class _TuneEntry(C.Structure):
    _fields_ = [("instances", C.c_uint64), ("max_target", C.c_uint64), ("target_digests", C.c_uint32),
                ("method", C.c_int32), ("chunk", C.c_uint32), ("ms", C.c_float)]


class _Info(C.Structure):
    _fields_ = [("name", C.c_char_p), ("orig_size", C.c_uint64), ("bf_count", C.c_uint32),
                ("last_count", C.c_uint32), ("dd_block_size", C.c_uint32), ("max_elem", C.c_uint32)]


def load():
    """Loads libsdbfx.so (built by sdhash_amd.build); raises when absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SdbfxError(f"{LIB_PATH} is not built (run python -m sdhash_amd.build)")
    L = C.CDLL(LIB_PATH)
    P, U32, U64, I32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32
    L.sdbfx_ctx_create.argtypes = [C.c_int, U64, C.POINTER(P)]
    L.sdbfx_ctx_destroy.argtypes = [P]
    L.sdbfx_ctx_destroy.restype = None
    L.sdbfx_last_error.argtypes = [P]
    L.sdbfx_last_error.restype = C.c_char_p
    L.sdbfx_sync.argtypes = [P]
    L.sdbfx_ctx_scratch_bytes.argtypes = [P]
    L.sdbfx_ctx_scratch_bytes.restype = U64
    L.sdbfx_ctx_trim.argtypes = [P]
    L.sdbfx_set_create.argtypes = [P, C.POINTER(P)]
    L.sdbfx_set_destroy.argtypes = [P]
    L.sdbfx_set_destroy.restype = None
    L.sdbfx_set_size.argtypes = [P]
    L.sdbfx_set_size.restype = U32
    L.sdbfx_set_filter_count.argtypes = [P]
    L.sdbfx_set_filter_count.restype = U64
    L.sdbfx_set_digest.argtypes = [P, P, U32, P, P, P, P, P]
    L.sdbfx_set_digest_device.argtypes = [P, P, U32, P, P, P, P, P, P]
    L.sdbfx_set_add.argtypes = [P, P, C.c_char_p, U64, U32, U32, U32, U32, P, P]
    L.sdbfx_set_info.argtypes = [P, U32, C.POINTER(_Info)]
    L.sdbfx_set_read.argtypes = [P, P, U32, P, P, P]
    L.sdbfx_compare_all.argtypes = [P, P, P]
    L.sdbfx_compare_to.argtypes = [P, P, P, P]
    L.sdbfx_compare_to_sampled.argtypes = [P, P, P, U32, P]
    L.sdbfx_compare_to_range.argtypes = [P, P, U32, U32, P, P]
    L.sdbfx_device_alloc.argtypes = [P, U64, C.POINTER(P)]
    L.sdbfx_device_free.argtypes = [P, P]
    L.sdbfx_upload.argtypes = [P, P, P, U64]
    L.sdbfx_set_profiling.argtypes = [P, C.c_int]
    L.sdbfx_kernel_time.argtypes = [P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(U64)]
    L.sdbfx_reset_kernel_times.argtypes = [P]
    L.sdbfx_set_compare_method.argtypes = [P, C.c_int]
    L.sdbfx_digest_blocks_noalloc.argtypes = [P, P, U64, U32, P, P, P, C.POINTER(U32)]
    L.sdbfx_set_to_text.argtypes = [P, P, P, U64, C.POINTER(U64)]
    L.sdbfx_set_add_text.argtypes = [P, P, P, U64, C.POINTER(U32)]
    L.sdbfx_set_fold.argtypes = [P, P]
    L.sdbfx_set_folded.argtypes = [P]
    L.sdbfx_set_fold_digest.argtypes = [P, P, U32]
    L.sdbfx_set_digest_folded.argtypes = [P, U32]
    L.sdbfx_serve_create.argtypes = [C.c_int, P, U64, U32, U64, C.POINTER(P)]
    L.sdbfx_serve_destroy.argtypes = [P]
    L.sdbfx_serve_destroy.restype = None
    L.sdbfx_serve_error.argtypes = [P]
    L.sdbfx_serve_error.restype = C.c_char_p
    L.sdbfx_serve_ref_size.argtypes = [P]
    L.sdbfx_serve_ref_size.restype = U32
    L.sdbfx_serve_submit.argtypes = [P, P, U64, C.c_char_p, U32, C.POINTER(U64)]
    L.sdbfx_serve_poll.argtypes = [P, U64, P, C.POINTER(I32), C.POINTER(U64), C.POINTER(C.c_int)]
    L.sdbfx_serve_wait.argtypes = [P, U64, P, C.POINTER(I32), C.POINTER(U64)]
    L.sdbfx_serve_stats.argtypes = [P, C.POINTER(U64), C.POINTER(U64), C.POINTER(U64)]
    L.sdbfx_serve_admission.argtypes = [P, C.POINTER(U64), C.POINTER(U64)]
    L.sdbfx_graph_blocks_create.argtypes = [P, P, U64, U32, P, P, P, C.POINTER(U32), C.POINTER(P)]
    L.sdbfx_graph_launch.argtypes = [P, P]
    L.sdbfx_graph_destroy.argtypes = [P]
    L.sdbfx_graph_destroy.restype = None
    L.sdbfx_set_autotune.argtypes = [P, C.c_int]
    L.sdbfx_autotune_table.argtypes = [P, P, U32, C.POINTER(U32)]
    L.sdbfx_index_create.argtypes = [P, U64, C.POINTER(P)]
    L.sdbfx_index_destroy.argtypes = [P]
    L.sdbfx_index_destroy.restype = None
    L.sdbfx_index_size.argtypes = [P]
    L.sdbfx_index_size.restype = U64
    L.sdbfx_index_read.argtypes = [P, P, P, C.POINTER(U64)]
    L.sdbfx_index_write.argtypes = [P, P, P, U64]
    L.sdbfx_set_digest_indexed.argtypes = [P, P, U32, P, P, P, P, P, P, U32, P, U64, P]
    L.sdbfx_set_digest_device_indexed.argtypes = [P, P, U32, P, P, P, P, P, P, P, U32, P, U64, P]
    for name in SYMBOLS:
        getattr(L, name).restype = getattr(L, name).restype or I32
    for name in ("sdbfx_ctx_create", "sdbfx_sync", "sdbfx_ctx_trim", "sdbfx_set_create", "sdbfx_set_digest",
                 "sdbfx_set_digest_device", "sdbfx_set_add", "sdbfx_set_info", "sdbfx_set_read",
                 "sdbfx_set_to_text", "sdbfx_set_add_text", "sdbfx_set_fold", "sdbfx_set_fold_digest",
                 "sdbfx_compare_all", "sdbfx_compare_to", "sdbfx_compare_to_sampled", "sdbfx_compare_to_range", "sdbfx_device_alloc", "sdbfx_device_free",
                 "sdbfx_upload", "sdbfx_set_profiling", "sdbfx_kernel_time",
                 "sdbfx_reset_kernel_times", "sdbfx_set_compare_method", "sdbfx_digest_blocks_noalloc", "sdbfx_serve_create",
                 "sdbfx_serve_submit", "sdbfx_serve_poll", "sdbfx_serve_wait", "sdbfx_serve_stats",
                 "sdbfx_serve_admission", "sdbfx_index_create", "sdbfx_index_read", "sdbfx_index_write",
                 "sdbfx_set_digest_indexed", "sdbfx_set_digest_device_indexed", "sdbfx_set_autotune",
                 "sdbfx_autotune_table", "sdbfx_graph_blocks_create", "sdbfx_graph_launch"):
        getattr(L, name).restype = I32
    _lib = L
    return L


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def name_array(names) -> C.Array:
    """The digest calls' name list as the C array they pass to the library.
    Built once, it can be handed to many digest calls over the same inputs
    in place of a list (the array holds the name bytes)."""
    keep = [x if isinstance(x, bytes) else str(x).encode() for x in names]
    return (C.c_char_p * max(len(keep), 1))(*keep)


def _bytes_array(items) -> tuple[list, C.Array]:
    arrs = [np.ascontiguousarray(np.frombuffer(x, np.uint8) if isinstance(x, (bytes, bytearray))
                                 else x, dtype=np.uint8) for x in items]
    ptrs = (C.c_void_p * max(len(arrs), 1))(*[a.ctypes.data for a in arrs])
    return arrs, ptrs


@dataclass
class DigestInfo:
    name: str
    orig_size: int
    bf_count: int
    last_count: int
    dd_block_size: int
    max_elem: int


class DeviceBuffer:
    """Device memory with 256 readable padding bytes after `nbytes`."""

    def __init__(self, engine: "Engine", nbytes: int):
        self.engine = engine
        self.nbytes = nbytes
        p = C.c_void_p()
        engine._check(engine.lib.sdbfx_device_alloc(engine.ctx, nbytes, C.byref(p)))
        self.ptr = p

    def upload(self, offset: int, data: np.ndarray) -> None:
        a = np.ascontiguousarray(data, dtype=np.uint8)
        assert offset + a.size <= self.nbytes
        self.engine._check(self.engine.lib.sdbfx_upload(
            self.engine.ctx, C.c_void_p(self.ptr.value + offset), _ptr(a), a.size))

    def free(self) -> None:
        if self.ptr:
            self.engine.lib.sdbfx_device_free(self.engine.ctx, self.ptr)
            self.ptr = None


class DigestSet:
    """A device-resident collection of sdbf digests (the engine's sdbf_set)."""

    def __init__(self, engine: "Engine"):
        self.engine = engine
        p = C.c_void_p()
        engine._check(engine.lib.sdbfx_set_create(engine.ctx, C.byref(p)))
        self.handle = p

    def __len__(self) -> int:
        return int(self.engine.lib.sdbfx_set_size(self.handle))

    def filter_count(self) -> int:
        return int(self.engine.lib.sdbfx_set_filter_count(self.handle))

    def info(self, i: int) -> DigestInfo:
        x = _Info()
        if self.engine.lib.sdbfx_set_info(self.handle, i, C.byref(x)) != 0:
            raise IndexError(i)
        return DigestInfo(x.name.decode("utf-8", "surrogateescape"), x.orig_size, x.bf_count,
                          x.last_count, x.dd_block_size, x.max_elem)

    def read(self, i: int):
        """(filters uint8[bf_count*256], elem_counts uint16[bf_count], hamming uint16[bf_count])"""
        inf = self.info(i)
        f = np.zeros(inf.bf_count * 256, np.uint8)
        ec = np.zeros(inf.bf_count, np.uint16)
        hm = np.zeros(inf.bf_count, np.uint16)
        self.engine._check(self.engine.lib.sdbfx_set_read(self.engine.ctx, self.handle, i, _ptr(f),
                                                          _ptr(ec), _ptr(hm)))
        return f, ec, hm

    def to_text(self) -> bytes:
        """The sdbf text of every digest, in order (sdbf_set::to_string), with
        the base64 encoded on the device."""
        n = C.c_uint64(0)
        lib, ctx = self.engine.lib, self.engine.ctx
        self.engine._check(lib.sdbfx_set_to_text(ctx, self.handle, None, 0, C.byref(n)))
        out = np.empty(max(n.value, 1), np.uint8)
        self.engine._check(lib.sdbfx_set_to_text(ctx, self.handle, _ptr(out), n.value, C.byref(n)))
        return out[: n.value].tobytes()

    def fold(self) -> None:
        """Fast mode (sdbf::fast): folds every filter to 64 bytes in place on
        the device; folded sets compare with the fast cutoffs."""
        self.engine._check(self.engine.lib.sdbfx_set_fold(self.engine.ctx, self.handle))

    def fold_digest(self, i: int) -> None:
        """sdbf::fast on digest i only."""
        self.engine._check(self.engine.lib.sdbfx_set_fold_digest(self.engine.ctx, self.handle, i))

    @property
    def folded(self) -> bool:
        return self.engine.lib.sdbfx_set_folded(self.handle) == 1

    def digest_folded(self, i: int) -> bool:
        return self.engine.lib.sdbfx_set_digest_folded(self.handle, i) == 1

    def add_text(self, text: bytes) -> int:
        """Appends the digests of an sdbf text (sdbf_set's file reader: lines up
        to the first empty one), decoding the base64 on the device; returns
        how many were added.  Raises SdbfxError on a malformed line."""
        buf = np.frombuffer(bytes(text), np.uint8)
        n = C.c_uint32(0)
        self.engine._check(self.engine.lib.sdbfx_set_add_text(
            self.engine.ctx, self.handle, _ptr(buf) if buf.size else None, buf.size, C.byref(n)))
        return n.value

    def digest(self, bufs, names=None, dd_block=0) -> np.ndarray:
        """Digests host buffers into this set; returns the per-input status."""
        n = len(bufs)
        arrs, ptrs = _bytes_array(bufs)
        lens = np.array([a.size for a in arrs], np.uint64)
        dd = self._dd(dd_block, n)
        nm = self._names(names, n)
        st = np.zeros(max(n, 1), np.int32)
        self.engine._check(self.engine.lib.sdbfx_set_digest(
            self.engine.ctx, self.handle, n, ptrs, _ptr(lens), nm, _ptr(dd), _ptr(st)))
        return st[:n]

    def digest_indexed(self, bufs, names=None, dd_block=0, index=None, search=()):
        """digest() with an index_info: `index` (an Index, or None) gates
        stream-mode inputs and takes block-mode inserts; `search` (Indexes)
        are looked up by block-mode inputs instead.  Returns (status,
        matches uint32[blocks of the block-mode inputs, len(search)])."""
        n = len(bufs)
        arrs, ptrs = _bytes_array(bufs)
        lens = np.array([a.size for a in arrs], np.uint64)
        return self._indexed(n, lens, names, dd_block, index, search,
                             lambda *a: self.engine.lib.sdbfx_set_digest_indexed(
                                 self.engine.ctx, self.handle, n, ptrs, _ptr(lens), *a))

    def digest_device_indexed(self, dbuf: DeviceBuffer, offsets, lens, names=None, dd_block=0, index=None,
                              search=()):
        """digest_device() with an index_info, as digest_indexed."""
        n = len(lens)
        off = np.ascontiguousarray(offsets, np.uint64)
        ln = np.ascontiguousarray(lens, np.uint64)
        return self._indexed(n, ln, names, dd_block, index, search,
                             lambda *a: self.engine.lib.sdbfx_set_digest_device_indexed(
                                 self.engine.ctx, self.handle, n, dbuf.ptr, _ptr(off), _ptr(ln), *a))

    def _indexed(self, n, lens, names, dd_block, index, search, call):
        dd = self._dd(dd_block, n)
        nm = self._names(names, n)
        st = np.zeros(max(n, 1), np.int32)
        search = list(search)
        nblk = sum(int(l) // int(d) + (1 if int(l) % int(d) >= 512 else 0)
                   for l, d in zip(lens, dd[:n]) if d and l >= 512)
        m = np.zeros(max(nblk * len(search), 1), np.uint32)
        sarr = (C.c_void_p * max(len(search), 1))(*[x.handle for x in search])
        self.engine._check(call(nm, _ptr(dd), index.handle if index else None, sarr, len(search), _ptr(m), nblk,
                                _ptr(st)))
        return st[:n], m[:nblk * len(search)].reshape(nblk, len(search))

    def digest_device(self, dbuf: DeviceBuffer, offsets, lens, names=None, dd_block=0) -> np.ndarray:
        n = len(lens)
        off = np.ascontiguousarray(offsets, np.uint64)
        ln = np.ascontiguousarray(lens, np.uint64)
        dd = self._dd(dd_block, n)
        nm = self._names(names, n)
        st = np.zeros(max(n, 1), np.int32)
        self.engine._check(self.engine.lib.sdbfx_set_digest_device(
            self.engine.ctx, self.handle, n, dbuf.ptr, _ptr(off), _ptr(ln), nm, _ptr(dd), _ptr(st)))
        return st[:n]

    def add(self, name: bytes, orig_size: int, bf_count: int, last_count: int, dd_block_size: int,
            max_elem: int, filters: np.ndarray, elem_counts: np.ndarray | None) -> None:
        f = np.ascontiguousarray(filters, np.uint8)
        ec = None if elem_counts is None else np.ascontiguousarray(elem_counts, np.uint16)
        self.engine._check(self.engine.lib.sdbfx_set_add(
            self.engine.ctx, self.handle, name, orig_size, bf_count, last_count, dd_block_size,
            max_elem, _ptr(f), None if ec is None else _ptr(ec)))

    @staticmethod
    def _dd(dd_block, n):
        if np.isscalar(dd_block):
            return np.full(max(n, 1), int(dd_block), np.uint32)
        return np.ascontiguousarray(dd_block, np.uint32)

    def _names(self, names, n):
        if isinstance(names, C.Array):   # from name_array(): passed as is
            if len(names) < n:
                raise ValueError("name array holds %d names for %d inputs" % (len(names), n))
            self._keep = names
            return names
        if names is None:
            names = [b"f%06d" % i for i in range(n)]
        self._keep = [x if isinstance(x, bytes) else str(x).encode() for x in names]
        return (C.c_char_p * max(n, 1))(*self._keep)

    def close(self) -> None:
        if self.handle:
            self.engine.lib.sdbfx_set_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Index:
    """A device-resident index (sdhash --index): bloom_filter(size, 5, 0,
    0.01) bits and element count."""

    def __init__(self, engine: "Engine", size: int = 64 << 20):
        self.engine = engine
        p = C.c_void_p()
        engine._check(engine.lib.sdbfx_index_create(engine.ctx, size, C.byref(p)))
        self.handle = p
        self.size = size

    def read(self) -> tuple[np.ndarray, int]:
        """(bits as uint8[size], element count)"""
        bits = np.zeros(self.size, np.uint8)
        e = C.c_uint64()
        self.engine._check(self.engine.lib.sdbfx_index_read(self.engine.ctx, self.handle, _ptr(bits), C.byref(e)))
        return bits, e.value

    def write(self, bits, elem_count: int) -> None:
        b = np.ascontiguousarray(bits, np.uint8)
        assert b.size == self.size
        self.engine._check(self.engine.lib.sdbfx_index_write(self.engine.ctx, self.handle, _ptr(b), elem_count))

    def close(self) -> None:
        if self.handle:
            self.engine.lib.sdbfx_index_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """One device context (HIP stream) of the engine."""

    def __init__(self, device: int = 0, mem_budget: int = 0):
        self.lib = load()
        p = C.c_void_p()
        if self.lib.sdbfx_ctx_create(device, mem_budget, C.byref(p)) != 0:
            raise SdbfxError(f"cannot open HIP device {device}")
        self.ctx = p

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise SdbfxError(self.lib.sdbfx_last_error(self.ctx).decode(), rc)

    def scratch_bytes(self) -> int:
        return int(self.lib.sdbfx_ctx_scratch_bytes(self.ctx))

    def trim(self) -> None:
        """Frees the context's scratch memory."""
        self._check(self.lib.sdbfx_ctx_trim(self.ctx))

    def new_set(self) -> DigestSet:
        return DigestSet(self)

    def new_index(self, size: int = 64 << 20) -> "Index":
        return Index(self, size)

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def compare_all(self, s: DigestSet) -> np.ndarray:
        n = len(s)
        out = np.zeros(max(n * (n - 1) // 2, 1), np.int32)
        self._check(self.lib.sdbfx_compare_all(self.ctx, s.handle, _ptr(out)))
        return out[: n * (n - 1) // 2]

    def compare_to(self, q: DigestSet, t: DigestSet, sample: int = 0) -> np.ndarray:
        """scores[i, j] = q[i].compare(t[j], sample) (sdhash -c, -s sample)."""
        out = np.zeros((max(len(q), 1), max(len(t), 1)), np.int32)
        if sample:
            self._check(self.lib.sdbfx_compare_to_sampled(self.ctx, q.handle, t.handle, sample, _ptr(out)))
        else:
            self._check(self.lib.sdbfx_compare_to(self.ctx, q.handle, t.handle, _ptr(out)))
        return out[: len(q), : len(t)]

    def sync(self) -> None:
        self._check(self.lib.sdbfx_sync(self.ctx))

    def compare_method(self, method: int) -> None:
        """CMP_AUTO, CMP_VALU (popcount on the vector ALUs), CMP_MFMA (matrix
        cores, filter bits as FP4 elements) or CMP_MFMA_I8 (matrix cores,
        int8 {0,1} byte planes); the scores are the same."""
        self._check(self.lib.sdbfx_set_compare_method(self.ctx, method))

    def autotune(self, on: bool) -> None:
        """Times every comparison configuration on the first call of each
        shape class and keeps the fastest (scores unchanged)."""
        self._check(self.lib.sdbfx_set_autotune(self.ctx, 1 if on else 0))

    def tuned(self) -> list[dict]:
        """The autotuned classes: instances, largest target, target digests,
        method (CMP_VALU / CMP_MFMA), chunk, ms."""
        n = C.c_uint32()
        self._check(self.lib.sdbfx_autotune_table(self.ctx, None, 0, C.byref(n)))
        arr = (_TuneEntry * max(n.value, 1))()
        self._check(self.lib.sdbfx_autotune_table(self.ctx, arr, n.value, C.byref(n)))
        return [{"instances": e.instances, "max_target": e.max_target, "target_digests": e.target_digests,
                 "method": "mfma" if e.method == CMP_MFMA else "valu", "chunk": e.chunk, "ms": e.ms}
                for e in arr[:n.value]]

    def profiling(self, on: bool) -> None:
        self._check(self.lib.sdbfx_set_profiling(self.ctx, 1 if on else 0))

    def kernel_time(self, name: str) -> tuple[float, int]:
        ms = C.c_double()
        n = C.c_uint64()
        self._check(self.lib.sdbfx_kernel_time(self.ctx, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def reset_kernel_times(self) -> None:
        self._check(self.lib.sdbfx_reset_kernel_times(self.ctx))

    def close(self) -> None:
        if self.ctx:
            self.lib.sdbfx_ctx_destroy(self.ctx)
            self.ctx = None


@dataclass
class ServeResult:
    status: int                 # 0, -3 (query under 512 bytes) or -1 (device error)
    scores: np.ndarray          # int32 per reference digest (empty unless status == 0)
    latency_s: float            # submit to result


class Server:
    """Continuous-batching query server over a resident reference set
    (sdbfx_serve_*): submit() returns a ticket at once, wait()/poll() give
    the query's scores against every reference digest."""

    def __init__(self, ref: DigestSet, device: int = 0, batch_bytes: int = 256 << 20,
                 batch_max: int = 1024, mem_budget: int = 0):
        self.lib = load()
        self.ref = ref                      # must outlive the server
        p = C.c_void_p()
        if self.lib.sdbfx_serve_create(device, ref.handle, batch_bytes, batch_max, mem_budget, C.byref(p)) != 0:
            raise SdbfxError("cannot create the server")
        self.handle = p
        self.nref = int(self.lib.sdbfx_serve_ref_size(p))

    def submit(self, data, name: bytes = b"query", dd_block: int = 0) -> int:
        a = np.ascontiguousarray(np.frombuffer(data, np.uint8) if isinstance(data, (bytes, bytearray))
                                 else data, dtype=np.uint8)
        t = C.c_uint64()
        if self.lib.sdbfx_serve_submit(self.handle, _ptr(a), a.size, name, dd_block, C.byref(t)) != 0:
            raise SdbfxError(self.lib.sdbfx_serve_error(self.handle).decode() or "submit failed")
        return t.value

    def _result(self, fn, ticket, *extra):
        sc = np.zeros(max(self.nref, 1), np.int32)
        st = C.c_int32()
        lat = C.c_uint64()
        rc = fn(self.handle, ticket, _ptr(sc), C.byref(st), C.byref(lat), *extra)
        if rc != 0:
            raise SdbfxError("unknown ticket %d" % ticket)
        return ServeResult(st.value, sc[: self.nref] if st.value == 0 else sc[:0], lat.value / 1e9)

    def wait(self, ticket: int) -> ServeResult:
        return self._result(self.lib.sdbfx_serve_wait, ticket)

    def poll(self, ticket: int):
        done = C.c_int()
        r = self._result(self.lib.sdbfx_serve_poll, ticket, C.byref(done))
        return r if done.value else None

    def stats(self) -> tuple[int, int]:
        """(batches, queries) completed so far."""
        return self.stats3()[:2]

    def stats3(self) -> tuple[int, int, int]:
        """(batches, queries, batches compared with the MFMA kernel)."""
        b, q, m = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.lib.sdbfx_serve_stats(self.handle, C.byref(b), C.byref(q), C.byref(m))
        return b.value, q.value, m.value

    def admission(self) -> tuple[int, int]:
        """(queries deferred to a later batch, queries refused) by the memory budget."""
        d, r = C.c_uint64(), C.c_uint64()
        self.lib.sdbfx_serve_admission(self.handle, C.byref(d), C.byref(r))
        return d.value, r.value

    def error(self) -> str:
        return self.lib.sdbfx_serve_error(self.handle).decode()

    def close(self) -> None:
        if self.handle:
            self.lib.sdbfx_serve_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

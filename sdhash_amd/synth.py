"""Deterministic synthetic inputs of the "segment mix" used by the tests and
the benchmark.

Every input is a concatenation of 4 KiB segments, drawn
  40 % uniform random bytes,
  40 % Zipf(s=1.1) over a 64-symbol alphabet (integer CDF table),
  10 % a random 16-byte motif repeated,
  10 % zeros,
and sets of files get chunks of 4-64 KiB copied between them ("planted"),
so that pairwise similarity scores span 0-100.  All draws come from numpy's
PCG64 integer streams, so a seed gives the same bytes on every host.
"""
from __future__ import annotations

import os

import numpy as np

SEG = 4096
KIB = 1024
MIB = 1 << 20

# Zipf(s=1.1) over 64 ranks as an integer CDF with total 2**16, and its
# inverse as a 65536-entry lookup table (rank of a 16-bit uniform draw).
_W = np.arange(1, 65, dtype=np.float64) ** -1.1
_CDF = np.round(np.cumsum(_W) / _W.sum() * (1 << 16)).astype(np.int64)
_CDF[-1] = 1 << 16
_INV = np.searchsorted(_CDF, np.arange(1 << 16), side="right").astype(np.uint8)


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def segment_mix(rng: np.random.Generator, nbytes: int) -> np.ndarray:
    """nbytes of segment-mix data (the last segment is cut to size)."""
    nseg = (nbytes + SEG - 1) // SEG
    out = np.zeros((nseg, SEG), dtype=np.uint8)
    kind = rng.integers(0, 10, nseg)
    uni = np.flatnonzero(kind < 4)
    zpf = np.flatnonzero((kind >= 4) & (kind < 8))
    mot = np.flatnonzero(kind == 8)
    if uni.size:
        out[uni] = rng.integers(0, 256, (uni.size, SEG), dtype=np.uint8)
    if zpf.size:
        keys = rng.integers(0, 1 << 32, (zpf.size, 256), dtype=np.uint64)
        alpha = np.argsort(keys, axis=1, kind="stable")[:, :64].astype(np.uint8)
        draw = rng.integers(0, 1 << 16, (zpf.size, SEG), dtype=np.uint16)
        sym = _INV[draw]
        out[zpf] = alpha[np.arange(zpf.size)[:, None], sym]
    if mot.size:
        motif = rng.integers(0, 256, (mot.size, 16), dtype=np.uint8)
        out[mot] = np.tile(motif, (1, SEG // 16))
    return out.reshape(-1)[:nbytes].copy()


def plant(rng: np.random.Generator, files: list[np.ndarray], max_frac: float,
          lo: int = 4 * KIB, hi: int = 64 * KIB) -> list[np.ndarray]:
    """Copy 4-64 KiB chunks of one partner file's original content into each
    file until a fraction drawn from [0, max_frac] of it is covered."""
    src = [f.copy() for f in files]
    out = [f.copy() for f in files]
    n = len(files)
    if n < 2:
        return out
    for i in range(n):
        size = out[i].size
        frac = int(rng.integers(0, int(round(max_frac * 100)) + 1)) / 100.0
        target = int(frac * size)
        # one partner file per receiver, so that planted pairs share a lot
        s = int(rng.integers(0, n - 1))
        s = s + 1 if s >= i else s
        done = 0
        while done < target:
            c = int(rng.integers(lo, hi + 1))
            if src[s].size < c or size < c:
                break
            so = int(rng.integers(0, src[s].size - c + 1))
            do = int(rng.integers(0, size - c + 1))
            out[i][do:do + c] = src[s][so:so + c]
            done += c
    return out


def file_set(seed: int, n: int, size: int, max_frac: float = 0.5) -> list[np.ndarray]:
    """n files of `size` bytes with planted cross-file chunks (0..max_frac)."""
    rng = _rng(seed)
    files = [segment_mix(rng, size) for _ in range(n)]
    return plant(rng, files, max_frac)


def ragged_set(seed: int, sizes, max_frac: float = 0.05) -> list[np.ndarray]:
    """Files of the given sizes with planted chunks."""
    rng = _rng(seed)
    files = [segment_mix(rng, int(s)) for s in sizes]
    return plant(rng, files, max_frac)


def lognormal_sizes(seed: int, n: int, median: int = 256 * KIB, sigma: float = 1.5,
                    lo: int = 4 * KIB, hi: int = 64 * MIB) -> np.ndarray:
    rng = _rng(seed)
    z = rng.standard_normal(n)
    s = np.exp(np.log(median) + sigma * z)
    return np.clip(s, lo, hi).astype(np.int64)


def disk_image(seed: int, nbytes: int, block: int = 64 * MIB) -> np.ndarray:
    """A disk image of segment-mix data, generated in pieces."""
    rng = _rng(seed)
    out = np.empty(nbytes, dtype=np.uint8)
    pos = 0
    while pos < nbytes:
        n = min(block, nbytes - pos)
        out[pos:pos + n] = segment_mix(rng, n)
        pos += n
    return out


def fragments(seed: int, image: np.ndarray, n: int, size: int = 4 * KIB) -> list[np.ndarray]:
    """n fragments: the first half cut from the image at random offsets,
    the second half fresh uniform random bytes."""
    rng = _rng(seed)
    half = n // 2
    offs = rng.integers(0, image.size - size + 1, half)
    out = [image[o:o + size].copy() for o in offs]
    fresh = rng.integers(0, 256, (n - half, size), dtype=np.uint8)
    out.extend(fresh[i] for i in range(n - half))
    return out


_synth_lib = None


def _synth():
    import ctypes as C
    global _synth_lib
    if _synth_lib is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsdbfx_synth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (run python -m sdhash_amd.build)")
        L = C.CDLL(path)
        L.sdbfx_synth_fill.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64]
        L.sdbfx_synth_fill.restype = C.c_int
        L.sdbfx_synth_plant.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.sdbfx_synth_plant.restype = C.c_int
        _synth_lib = L
    return _synth_lib


def device_fill(ptr: int, nbytes: int, seed: int, offset: int = 0) -> None:
    """Fills device memory at `ptr` with `nbytes` of the device-side segment
    mix (libsdbfx_synth.so, same mix as segment_mix, different bytes), the
    stream `seed` from byte `offset` on.  nbytes, offset: multiples of 16."""
    import ctypes as C
    if _synth().sdbfx_synth_fill(C.c_void_p(ptr), nbytes, offset, seed) != 0:
        raise RuntimeError("device fill failed")


def device_plant(dst: int, dst_size: int, src: int, src_size: int, d_off, s_off, lens) -> None:
    """Copies chunks (offsets and lengths multiples of 16) from device memory
    src + s_off[i] to dst + d_off[i] in one launch (libsdbfx_synth.so); src
    is not written.  Every chunk must lie inside both buffers."""
    import ctypes as C
    d = np.ascontiguousarray(d_off, np.uint64)
    s = np.ascontiguousarray(s_off, np.uint64)
    n = np.ascontiguousarray(lens, np.uint32)
    if not (d.size == s.size == n.size):
        raise ValueError("offset and length arrays differ in size")
    if n.size and (int((d + n).max()) > dst_size or int((s + n).max()) > src_size):
        raise ValueError("a planted chunk lies outside its buffer")
    if n.size and _synth().sdbfx_synth_plant(C.c_void_p(dst), C.c_void_p(src), n.size, d.ctypes.data,
                                             s.ctypes.data, n.ctypes.data) != 0:
        raise RuntimeError("device plant failed")

"""The no-allocation block digest (sdbfx_digest_blocks_noalloc) through the C
ABI on the GPU: its filters, element counts and Hamming weights equal those
of sdbfx_set_digest_device for the same input, on two calls in a row (the
second with a smaller input, over the scratch array the first one left);
sizes beyond its maximum and invalid arguments are refused."""
import ctypes as C

import numpy as np
import pytest

import oracle
import sdhash_amd
from sdhash_amd import synth

pytestmark = pytest.mark.gpu

KIB = 1024
MIB = 1 << 20


def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemcpy.restype = C.c_int
    return h


def download(h, ptr, nbytes, dtype):
    out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
    assert h.hipMemcpy(out.ctypes.data, ptr, nbytes, 2) == 0
    return out


def noalloc(eng, dbuf, size, dd, fbuf, ebuf, hbuf):
    n = C.c_uint32()
    rc = eng.lib.sdbfx_digest_blocks_noalloc(eng.ctx, dbuf.ptr, size, dd, fbuf.ptr, ebuf.ptr, hbuf.ptr, C.byref(n))
    return rc, n.value


def test_noalloc_block_digest_matches_set_digest():
    eng = sdhash_amd.Engine(0)
    h = hip()
    big = 40 * MIB + 700
    nmax = big // 16384 + 1
    fbuf, ebuf, hbuf = eng.alloc(nmax * 256), eng.alloc(nmax * 2), eng.alloc(nmax * 2)
    img = synth.disk_image(401, big)
    for size, dd in ((big, 16384), (3 * MIB + 100, 4096), (5000, 1024)):
        d = eng.alloc(size)
        d.upload(0, img[:size])
        rc, n = noalloc(eng, d, size, dd, fbuf, ebuf, hbuf)
        assert rc == 0, eng.lib.sdbfx_last_error(eng.ctx)
        eng.sync()
        assert n == size // dd + (1 if size % dd >= 512 else 0)
        f = download(h, fbuf.ptr.value, n * 256, np.uint8)
        ec = download(h, ebuf.ptr.value, n * 2, np.uint16)
        hm = download(h, hbuf.ptr.value, n * 2, np.uint16)
        s = eng.new_set()
        assert np.all(s.digest_device(d, [0], [size], [b"x"], dd_block=dd) == 0)
        wf, wec, whm = s.read(0)
        assert np.array_equal(f, wf) and np.array_equal(ec, wec) and np.array_equal(hm, whm), (size, dd)
        od = oracle.digest_block(img[:size], dd, b"x", threads=oracle.threads_available())
        assert np.array_equal(f, od.filters) and np.array_equal(ec, od.elem_counts), (size, dd)
        s.close()
        d.free()
    # refused: too many blocks, a bad dd-block size, a too-small input
    d = eng.alloc(2 * MIB)
    rc, n = noalloc(eng, d, 131073 * 512, 512, fbuf, ebuf, hbuf)
    assert rc == -1 and b"at most" in eng.lib.sdbfx_last_error(eng.ctx)
    rc, _ = noalloc(eng, d, MIB, 1000, fbuf, ebuf, hbuf)
    assert rc == -1
    rc, _ = noalloc(eng, d, 500, 16384, fbuf, ebuf, hbuf)
    assert rc == -3
    eng.close()

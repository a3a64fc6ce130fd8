"""The no-allocation block digest captured into a HIP graph
(sdbfx_graph_blocks_create / sdbfx_graph_launch) on the GPU: the buffer is
refilled eagerly between replays, with other eager work and synchronisations
in between, and every replay's filters, element counts and Hamming weights
must equal an eager digest of the same bytes."""
import ctypes as C

import numpy as np
import pytest

import oracle
import sdhash_amd
from sdhash_amd import synth

from test_zz_no_alloc import download, hip

pytestmark = pytest.mark.gpu

MIB = 1 << 20


def test_graph_replays_match_eager(engine):
    h = hip()
    size, dd = 3 * MIB + 700, 16384
    n = size // dd + (1 if size % dd >= 512 else 0)
    d = engine.alloc(size)
    fbuf, ebuf, hbuf = engine.alloc(n * 256), engine.alloc(n * 2), engine.alloc(n * 2)
    g = C.c_void_p()
    nf = C.c_uint32()
    rc = engine.lib.sdbfx_graph_blocks_create(engine.ctx, d.ptr, size, dd, fbuf.ptr, ebuf.ptr, hbuf.ptr,
                                              C.byref(nf), C.byref(g))
    assert rc == 0, engine.lib.sdbfx_last_error(engine.ctx)
    assert nf.value == n
    try:
        for k in range(4):
            data = synth.disk_image(610 + k, size)
            d.upload(0, data)
            assert engine.lib.sdbfx_graph_launch(engine.ctx, g) == 0
            engine.sync()
            f = download(h, fbuf.ptr.value, n * 256, np.uint8)
            ec = download(h, ebuf.ptr.value, n * 2, np.uint16)
            hm = download(h, hbuf.ptr.value, n * 2, np.uint16)
            # eager work between replays: the set digest of the same bytes
            s = engine.new_set()
            assert np.all(s.digest_device(d, [0], [size], [b"x"], dd_block=dd) == 0)
            wf, wec, whm = s.read(0)
            s.close()
            assert np.array_equal(f, wf) and np.array_equal(ec, wec) and np.array_equal(hm, whm), k
            if k == 0:
                od = oracle.digest_block(data, dd, b"x", threads=oracle.threads_available())
                assert np.array_equal(f, od.filters) and np.array_equal(ec, od.elem_counts)
    finally:
        engine.lib.sdbfx_graph_destroy(g)
    # the checks of the no-allocation entry point
    assert engine.lib.sdbfx_graph_blocks_create(engine.ctx, d.ptr, 500, dd, fbuf.ptr, ebuf.ptr, hbuf.ptr,
                                                C.byref(nf), C.byref(g)) == -3
    assert engine.lib.sdbfx_graph_blocks_create(engine.ctx, d.ptr, size, 1000, fbuf.ptr, ebuf.ptr, hbuf.ptr,
                                                C.byref(nf), C.byref(g)) == -1

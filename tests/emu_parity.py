"""Parity cases run against the CPU emulation of the kernels (see
tests/hipemu): the unmodified kernel source, built for the host with
AddressSanitizer, digests and compares small inputs that are checked
against the oracle.  Run by tests/test_kernels_emulated.py in a subprocess
with libasan preloaded:  python tests/emu_parity.py <libsdbfx_emu.so>"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
import sdhash_amd  # noqa: E402
from sdhash_amd import synth  # noqa: E402

KIB = 1024


def same(gset, i, od, label):
    f, ec, hm = gset.read(i)
    info = gset.info(i)
    assert info.bf_count == od.bf_count, (label, info.bf_count, od.bf_count)
    assert np.array_equal(f, od.filters), label
    od.cstruct()
    assert np.array_equal(hm, od.hamming[: od.bf_count]), label
    if od.elem_counts is not None:
        assert np.array_equal(ec, od.elem_counts), label
    else:
        assert info.last_count == od.last_count, label


def main(lib):
    sdhash_amd.LIB_PATH = lib
    eng = sdhash_amd.Engine(0)
    rng = np.random.Generator(np.random.PCG64(5))

    # block mode: full blocks, tails >= 512 and < 512, tiny inputs, dd 4096
    img = synth.disk_image(6, 3 * 16 * KIB + 700)
    pieces = [img, img[: 2 * 16 * KIB + 300], img[:512], img[:513], img[: 16 * KIB - 1]]
    s = eng.new_set()
    assert np.all(s.digest(pieces, dd_block=16384) == 0)
    for i, p in enumerate(pieces):
        same(s, i, oracle.digest_block(p, 16384), "block %d" % i)
    s = eng.new_set()
    assert np.all(s.digest([img], dd_block=4096) == 0)
    same(s, 0, oracle.digest_block(img, 4096), "dd 4096")
    print("block ok", flush=True)

    # stream mode: one tile, tile boundaries, several tiles, degenerate data
    # (the last one passes 8738 de-duplication elements: a reset)
    sizes = [512, 600, 4 * KIB, 16 * KIB + 64, 16 * KIB + 65, 40000, 70000, 900 * KIB]
    bufs = [synth.segment_mix(rng, n) for n in sizes]
    bufs += [np.zeros(20000, np.uint8), np.full(20000, 7, np.uint8),
             np.tile(np.arange(64, dtype=np.uint8), 400), np.tile(np.arange(16, dtype=np.uint8), 2000)]
    s = eng.new_set()
    assert np.all(s.digest(bufs) == 0)
    for i, b in enumerate(bufs):
        same(s, i, oracle.digest_stream(b), "stream %d (%d bytes)" % (i, b.size))
    st = eng.new_set().digest([np.zeros(100, np.uint8), np.ones(512, np.uint8)])
    assert list(st) == [-3, 0]
    print("stream ok", flush=True)

    # compare: all pairs of mixed-mode digests, orientation by name
    files = synth.file_set(7, 8, 40 * KIB)
    names = [b"z", b"a", b"m", b"b", b"y", b"c", b"x", b"d"]
    modes = [0, 16384, 0, 0, 16384, 0, 0, 16384]
    s = eng.new_set()
    assert np.all(s.digest(files, names, dd_block=modes) == 0)
    ods = [oracle.digest_block(f, m, n) if m else oracle.digest_stream(f, n)
           for f, n, m in zip(files, names, modes)]
    for i, od in enumerate(ods):
        same(s, i, od, "set %d" % i)
    assert np.array_equal(eng.compare_all(s), oracle.compare_all(ods))
    # the text form of the set, base64 encoded by the device
    assert s.to_text() == "".join(od.to_string() for od in ods).encode()
    # ... and read back: headers on the host, bodies decoded by the device
    text = s.to_text()
    r = eng.new_set()
    assert r.add_text(text + b"\n" + b"junk after the empty line") == len(ods)
    for i, od in enumerate(ods):
        same(r, i, od, "text %d" % i)
        assert r.info(i).name.encode() == od.name and r.info(i).orig_size == od.orig_size
    assert r.to_text() == text
    assert np.array_equal(eng.compare_all(r), oracle.compare_all(ods))
    lines = text.split(b"\n")[:-1]
    bad = [lines[0][:300], b"sdbx" + lines[0][4:], lines[0][:-8], lines[1][:-9],
           lines[0].replace(b":256:", b":128:", 1), lines[0][:-5] + b"@" + lines[0][-4:],
           lines[1][:-5] + b"*" + lines[1][-4:]]
    for k, b in enumerate(bad):
        n0 = len(r)
        try:
            r.add_text(lines[2] + b"\n" + b + b"\n")
        except sdhash_amd.SdbfxError as e:
            assert "text" in str(e), (k, str(e))
        else:
            raise AssertionError("malformed line %d accepted" % k)
        assert len(r) == n0
    assert r.add_text(b"") == 0 and r.add_text(b"\n" + lines[0]) == 0
    print("text ok", flush=True)
    # compare_to with a target digest of more than 1024 filters (chunked)
    big = synth.disk_image(8, 1100 * 16 * KIB)
    t = [oracle.digest_block(big, 16384, b"big")]
    q = [oracle.digest_stream(big[o:o + 20000], b"q%d" % k) for k, o in enumerate((0, 5000000, 17000000))]
    q.append(oracle.digest_stream(files[0], b"other"))
    gt, gq = eng.new_set(), eng.new_set()
    for d in t:
        gt.add(d.name, d.orig_size, d.bf_count, d.last_count, d.dd_block_size, d.max_elem, d.filters,
               d.elem_counts)
    for d in q:
        gq.add(d.name, d.orig_size, d.bf_count, d.last_count, d.dd_block_size, d.max_elem, d.filters,
               d.elem_counts)
    want = oracle.compare_to(q, t)
    assert np.array_equal(eng.compare_to(gq, gt), want), (eng.compare_to(gq, gt), want)
    assert np.array_equal(eng.compare_to(gt, gq), oracle.compare_to(t, q))
    assert (want > 0).sum() >= 3
    # more than one 128-instance row block against a two-chunk target
    q2 = [oracle.digest_stream(big[o:o + 30000], b"r%d" % k)
          for k, o in enumerate(range(0, 17000000, 170000))]
    gq2 = eng.new_set()
    for d in q2:
        gq2.add(d.name, d.orig_size, d.bf_count, d.last_count, d.dd_block_size, d.max_elem, d.filters,
                d.elem_counts)
    assert sum(d.bf_count for d in q2) > 128
    want2 = oracle.compare_to(q2, t)
    assert (want2 > 0).sum() >= 10
    want_all = oracle.compare_all(ods)
    # identical digests: a candidate in every row of a tile (more than 64)
    d = oracle.digest_stream(synth.file_set(9, 1, 400 * KIB)[0], b"x")
    dups = [oracle.Digest(b"d%d" % i, d.filters, d.bf_count, d.last_count, None, d.orig_size, 0)
            for i in range(4)]
    gd = eng.new_set()
    for x in dups:
        gd.add(x.name, x.orig_size, x.bf_count, x.last_count, 0, 160, x.filters, None)
    want_dup = oracle.compare_all(dups)
    assert d.bf_count > 20 and np.all(want_dup == 100)
    # one filter repeated: every pair of a tile is a candidate (listing overflows)
    one = d.filters[:256]
    reps = [oracle.Digest(b"r%d" % i, np.tile(one, n), n, 160, None, n * 40000, 0) for i, n in enumerate((150, 140))]
    gr = eng.new_set()
    for x in reps:
        gr.add(x.name, x.orig_size, x.bf_count, x.last_count, 0, 160, x.filters, None)
    want_rep = oracle.compare_all(reps)
    for method in (sdhash_amd.CMP_MFMA, sdhash_amd.CMP_MFMA_I8, sdhash_amd.CMP_VALU):
        eng.compare_method(method)
        assert np.array_equal(eng.compare_all(s), want_all), method
        assert np.array_equal(eng.compare_to(gq, gt), want), method
        assert np.array_equal(eng.compare_to(gq2, gt), want2), method
        assert np.array_equal(eng.compare_all(gd), want_dup), method
        assert np.array_equal(eng.compare_all(gr), want_rep), method
        print("compare method %d ok" % method, flush=True)
    # a target set that grows after the column order of its first digest
    # was built: the new digest is ordered on its own, the old order kept
    eng.compare_method(sdhash_amd.CMP_MFMA)
    t2 = oracle.digest_block(big[: 300 * 16 * KIB], 4096, b"big2")
    gt.add(t2.name, t2.orig_size, t2.bf_count, t2.last_count, t2.dd_block_size, t2.max_elem, t2.filters,
           t2.elem_counts)
    assert np.array_equal(eng.compare_to(gq2, gt), oracle.compare_to(q2, t + [t2]))
    eng.compare_method(sdhash_amd.CMP_AUTO)
    # fast mode (sdbf::fast): folded filters and Hamming weights, fast scores,
    # folding twice changes nothing
    fs = eng.new_set()
    assert np.all(fs.digest(files, names, dd_block=modes) == 0)
    fs.fold()
    fs.fold()
    assert fs.folded
    fods = [d.folded() for d in ods]
    for i, d in enumerate(fods):
        f, ec, hm = fs.read(i)
        assert np.array_equal(f, d.filters) and np.array_equal(hm, d.hamming[: d.bf_count]), i
    assert np.array_equal(eng.compare_all(fs), oracle.compare_all(fods))
    # folded and plain digests together: each pair scored by its reference
    # digest's mode (a plain reference against a folded target included)
    ft, fq = eng.new_set(), eng.new_set()
    for d, g in ((t[0], ft), *((x, fq) for x in q)):
        g.add(d.name, d.orig_size, d.bf_count, d.last_count, d.dd_block_size, d.max_elem, d.filters, d.elem_counts)
    ft.fold()
    assert np.array_equal(eng.compare_to(fq, ft), oracle.compare_to(q, [t[0].folded()]))
    assert np.array_equal(eng.compare_to(ft, fq), oracle.compare_to([t[0].folded()], q))
    fq.fold()
    assert np.array_equal(eng.compare_to(fq, ft), oracle.compare_to([x.folded() for x in q], [t[0].folded()]))
    mixed = eng.new_set()
    mods = []
    for i, d in enumerate(ods):
        mixed.add(d.name, d.orig_size, d.bf_count, d.last_count, d.dd_block_size, d.max_elem, d.filters,
                  d.elem_counts)
        if i % 2:
            mixed.fold_digest(i)
        mods.append(d.folded() if i % 2 else d)
    assert not mixed.folded and mixed.digest_folded(1) and not mixed.digest_folded(2)
    assert np.array_equal(eng.compare_all(mixed), oracle.compare_all(mods))
    # sampling (sdhash -s): a query digest's first `sample` filters
    for smp in (1, 2, 3, 7):
        assert np.array_equal(eng.compare_to(gq2, gt, sample=smp), oracle.compare_to(q2, t + [t2], sample=smp)), smp
        assert np.array_equal(eng.compare_to(gt, gq2, sample=smp), oracle.compare_to(t + [t2], q2, sample=smp)), smp
    print("fast ok", flush=True)
    print("compare ok", flush=True)

    # serving: queries batched with others and alone give the oracle's scores
    queries = [files[0][:30000], files[3], np.zeros(100, np.uint8), big[:20000], files[5][1000:21000]]
    qdd = [0, 16384, 0, 0, 0]
    want_q = [None if q.size < 512 else
              oracle.compare_to([oracle.digest_block(q, d, b"q") if d else oracle.digest_stream(q, b"q")], ods)[0]
              for q, d in zip(queries, qdd)]
    srv = sdhash_amd.Server(s, batch_bytes=64 * KIB, batch_max=2)
    tickets = [srv.submit(q, b"q%d" % i, d) for i, (q, d) in enumerate(zip(queries, qdd))]
    mixed = [srv.wait(t) for t in tickets]
    alone = []
    for i, (q, d) in enumerate(zip(queries, qdd)):
        alone.append(srv.wait(srv.submit(q, b"q%d" % i, d)))
    nb, nq = srv.stats()
    srv.close()
    assert nq == 2 * len(queries) and nb < nq, (nb, nq)
    for w, m, a in zip(want_q, mixed, alone):
        if w is None:
            assert m.status == -3 and a.status == -3
        else:
            assert m.status == 0 and np.array_equal(m.scores, w) and np.array_equal(a.scores, w), (m, a, w)
    print("serve ok", flush=True)

    # memory budget: a context whose scratch would exceed it refuses the
    # call (SDBFX_EBUDGET) and leaves the set as it was
    small = sdhash_amd.Engine(0, mem_budget=600 * KIB)
    bs = small.new_set()
    assert np.all(bs.digest([files[0][:30000]]) == 0)
    try:
        bs.digest([files[1], big[: 1200 * KIB]])
    except sdhash_amd.SdbfxError as e:
        assert e.status == sdhash_amd.EBUDGET, e.status
    else:
        raise AssertionError("budget not enforced")
    assert len(bs) == 1 and small.scratch_bytes() <= 600 * KIB
    small.trim()
    assert small.scratch_bytes() == 0
    assert np.all(bs.digest([files[1]]) == 0) and len(bs) == 2
    same(bs, 1, oracle.digest_stream(files[1]), "budget")
    small.close()
    # the server admits by memory: batches that do not fit are split and
    # their later queries deferred; a query that never fits is refused
    srv = sdhash_amd.Server(s, batch_bytes=16 << 20, batch_max=64, mem_budget=1200 * KIB)
    qs = [files[k % 8][: 24 * KIB + 512 * k] for k in range(12)] + [big[: 1200 * KIB]]
    want_s = [oracle.compare_to([oracle.digest_stream(x, b"q")], ods)[0] for x in qs[:-1]]
    tickets = [srv.submit(x, b"q%d" % i) for i, x in enumerate(qs)]
    res = [srv.wait(t) for t in tickets]
    deferred, refused = srv.admission()
    srv.close()
    for w, r in zip(want_s, res[:-1]):
        assert r.status == 0 and np.array_equal(r.scores, w), r
    assert res[-1].status == sdhash_amd.EBUDGET and refused == 1, (res[-1].status, refused)
    assert deferred > 0, deferred
    print("admission ok (%d deferred)" % deferred, flush=True)

    # the no-allocation block digest (work list in the fixed scratch array):
    # the same filters as the set digest, twice in a row, the second smaller
    import ctypes as C
    for size, dd in ((3 * 16 * KIB + 700, 16384), (9000, 1024)):
        d = eng.alloc(size)
        d.upload(0, img[:size] if size <= img.size else big[:size])
        n = size // dd + (1 if size % dd >= 512 else 0)
        fb, eb, hb = eng.alloc(n * 256), eng.alloc(n * 2), eng.alloc(n * 2)
        cnt = C.c_uint32()
        assert eng.lib.sdbfx_digest_blocks_noalloc(eng.ctx, d.ptr, size, dd, fb.ptr, eb.ptr, hb.ptr,
                                                   C.byref(cnt)) == 0
        eng.sync()
        assert cnt.value == n
        ns = eng.new_set()
        assert np.all(ns.digest_device(d, [0], [size], [b"x"], dd_block=dd) == 0)
        wf, wec, whm = ns.read(0)
        assert C.string_at(fb.ptr.value, n * 256) == wf.tobytes()
        assert C.string_at(eb.ptr.value, n * 2) == wec.tobytes()
        assert C.string_at(hb.ptr.value, n * 2) == whm.tobytes()
    # the same launches captured into a graph and replayed twice
    size, dd = 3 * 16 * KIB + 700, 16384
    d = eng.alloc(size)
    d.upload(0, img[:size])
    n = size // dd + 1
    fb, eb, hb = eng.alloc(n * 256), eng.alloc(n * 2), eng.alloc(n * 2)
    g = C.c_void_p()
    cnt = C.c_uint32()
    assert eng.lib.sdbfx_graph_blocks_create(eng.ctx, d.ptr, size, dd, fb.ptr, eb.ptr, hb.ptr, C.byref(cnt),
                                             C.byref(g)) == 0
    for _ in range(2):
        assert eng.lib.sdbfx_graph_launch(eng.ctx, g) == 0
        eng.sync()
        ns = eng.new_set()
        assert np.all(ns.digest_device(d, [0], [size], [b"x"], dd_block=dd) == 0)
        wf, _, _ = ns.read(0)
        assert C.string_at(fb.ptr.value, n * 256) == wf.tobytes()
    eng.lib.sdbfx_graph_destroy(g)
    print("noalloc ok", flush=True)

    # indexes (--index, --index-search): stream inputs gated by the index,
    # block inputs inserting, in the order given, sharing one small index
    # (64 KiB, so that positions collide); then block inputs searching it
    index_cases(eng, rng)
    print("index ok", flush=True)
    eng.close()


def index_cases(eng, rng):
    img = synth.disk_image(61, 200 * KIB)
    a = img[: 90 * KIB]
    b = np.concatenate([img[40 * KIB: 120 * KIB], synth.segment_mix(rng, 30 * KIB)])
    c = img[60 * KIB: 60 * KIB + 3 * 16 * KIB + 900]
    d = np.concatenate([img[70 * KIB: 110 * KIB], img[: 20 * KIB]])
    e = img[100 * KIB: 200 * KIB]
    bufs = [a, b, c, d, e, img[:300]]
    modes = [0, 0, 16384, 0, 4096, 0]
    ix = eng.new_index(64 * KIB)
    oix = oracle.Index(64 * KIB)
    s = eng.new_set()
    st, _ = s.digest_indexed(bufs, [b"i%d" % k for k in range(6)], dd_block=modes, index=ix)
    assert list(st) == [0, 0, 0, 0, 0, -3], st
    k = 0
    for buf, m in zip(bufs[:5], modes):
        od = oracle.digest_block_indexed(buf, m, index=oix)[0] if m else oracle.digest_stream_indexed(buf, oix)
        same(s, k, od, "indexed input %d" % k)
        k += 1
    bits, ecount = ix.read()
    assert np.array_equal(bits, oix.bf) and ecount == oix.elem_count, (ecount, oix.elem_count)
    # without the index the stream digests differ (the gate skipped features)
    plain = oracle.digest_stream(b)
    assert plain.bf_count != s.info(1).bf_count or not np.array_equal(plain.filters, s.read(1)[0])
    # search: the index above and a second one, loaded from the host
    other = oracle.Index(4 * KIB)
    for buf in (img[150 * KIB: 200 * KIB],):
        oracle.digest_block_indexed(buf, 16384, index=other)
    ix2 = eng.new_index(4 * KIB)
    ix2.write(other.bf, other.elem_count)
    sb = [img[: 100 * KIB], img[120 * KIB: 200 * KIB], synth.segment_mix(rng, 40 * KIB)]
    s2 = eng.new_set()
    st, got = s2.digest_indexed(sb, dd_block=[16384, 4096, 0], search=[ix, ix2])
    assert np.all(st == 0)
    want = [oracle.digest_block_indexed(x, m, search=[oix, other]) for x, m in zip(sb[:2], (16384, 4096))]
    assert np.array_equal(got, np.concatenate([w[1] for w in want])), (got, want)
    assert got.sum() > 0
    for k2, (w, _) in enumerate(want):
        same(s2, k2, w, "searching input %d" % k2)
    ix.close()
    ix2.close()


if __name__ == "__main__":
    main(sys.argv[1])

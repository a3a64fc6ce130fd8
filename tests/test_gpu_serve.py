"""The query server (sdbfx_serve_*) on the GPU: every query's scores against
the resident reference set equal the oracle's compare_to row, whether the
query was batched with others or served alone."""
import numpy as np
import pytest

import oracle
import sdhash_amd
from sdhash_amd import synth

pytestmark = pytest.mark.gpu

KIB = 1024
MIB = 1 << 20


def reference(engine):
    img = synth.disk_image(41, 8 * MIB)
    files = synth.file_set(42, 6, 300 * KIB)
    ods = [oracle.digest_block(img[:4 * MIB], 16384, b"img.0000M"),
           oracle.digest_block(img[4 * MIB:], 16384, b"img.0004M")]
    ods += [oracle.digest_stream(f, b"file%d" % i) for i, f in enumerate(files)]
    s = engine.new_set()
    for d in ods:
        s.add(d.name, d.orig_size, d.bf_count, d.last_count, d.dd_block_size, d.max_elem, d.filters,
              d.elem_counts)
    return img, files, ods, s


def queries(img, files):
    rng = np.random.Generator(np.random.PCG64(43))
    qs, dd = [], []
    for k in range(40):
        n = int(2 ** rng.uniform(12, 20))                       # 4 KiB .. 1 MiB, log-uniform
        src = img if k % 3 else files[k % len(files)]
        o = int(rng.integers(0, max(src.size - n, 1)))
        qs.append(src[o:o + n] if k % 5 else synth.segment_mix(rng, n))
        dd.append(16384 if (k % 7 == 0 and n >= 16384) else 0)
    qs.append(np.zeros(300, np.uint8))                          # too small: status -3
    dd.append(0)
    return qs, dd


def want_rows(qs, dd, ods):
    out = []
    for q, d in zip(qs, dd):
        if q.size < 512:
            out.append(None)
            continue
        od = oracle.digest_block(q, d, b"q") if d else oracle.digest_stream(q, b"q")
        out.append(oracle.compare_to([od], ods)[0])
    return out


def check(res, want):
    for r, w in zip(res, want):
        if w is None:
            assert r.status == -3
        else:
            assert r.status == 0 and np.array_equal(r.scores, w), (r, w)


def test_serve_batched_and_alone(engine):
    img, files, ods, ref = reference(engine)
    qs, dd = queries(img, files)
    want = want_rows(qs, dd, ods)
    assert sum(1 for w in want if w is not None and (w > 0).any()) > 10
    srv = sdhash_amd.Server(ref, batch_bytes=2 * MIB, batch_max=16)
    try:
        tickets = [srv.submit(q, b"q%03d" % i, d) for i, (q, d) in enumerate(zip(qs, dd))]
        mixed = [srv.wait(t) for t in tickets]
        nb_mixed, _ = srv.stats()
        alone = [srv.wait(srv.submit(q, b"q%03d" % i, d)) for i, (q, d) in enumerate(zip(qs, dd))]
        nb, nq = srv.stats()
        assert srv.error() == ""
    finally:
        srv.close()
    check(mixed, want)
    check(alone, want)
    assert nq == 2 * len(qs)
    assert nb_mixed < len(qs)                    # queries were batched
    assert nb - nb_mixed == len(qs)              # one batch per lone query


def test_serve_poll_and_unknown_ticket(engine):
    img, files, ods, ref = reference(engine)
    srv = sdhash_amd.Server(ref)
    try:
        t = srv.submit(img[:50000], b"p")
        r = None
        while r is None:
            r = srv.poll(t)
        want = oracle.compare_to([oracle.digest_stream(img[:50000], b"p")], ods)[0]
        assert r.status == 0 and np.array_equal(r.scores, want)
        with pytest.raises(sdhash_amd.SdbfxError):
            srv.wait(t)                          # released by poll
    finally:
        srv.close()


def test_serve_fast_mode(engine):
    # a folded reference set (sdhash --fast): the server folds each batch of
    # queries and scores them with the fast cutoffs
    img, files, ods, ref = reference(engine)
    ref.fold()
    fods = [d.folded() for d in ods]
    qs, dd = queries(img, files)
    want = []
    for q, d in zip(qs, dd):
        if q.size < 512:
            want.append(None)
            continue
        od = oracle.digest_block(q, d, b"q") if d else oracle.digest_stream(q, b"q")
        want.append(oracle.compare_to([od.folded()], fods)[0])
    assert sum(1 for w in want if w is not None and (w > 0).any()) > 5
    srv = sdhash_amd.Server(ref, batch_bytes=2 * MIB, batch_max=16)
    try:
        res = [srv.wait(t) for t in [srv.submit(q, b"q%03d" % i, d) for i, (q, d) in enumerate(zip(qs, dd))]]
        assert srv.error() == ""
    finally:
        srv.close()
    check(res, want)


def test_serve_admits_by_memory(engine):
    # a scratch budget that one 4 MiB query does not fit into: batches are
    # split (later queries deferred, still answered), the large query refused
    img, files, ods, ref = reference(engine)
    qs = [img[o:o + 100 * KIB] for o in range(0, 6 * MIB, 300 * KIB)] + [img[:4 * MIB]]
    want = [oracle.compare_to([oracle.digest_stream(q, b"q")], ods)[0] for q in qs[:-1]]
    srv = sdhash_amd.Server(ref, batch_bytes=64 * MIB, batch_max=64, mem_budget=6 * MIB)
    try:
        first = srv.submit(img[:50 * KIB], b"warm")   # keeps the digest stage busy while the rest queue up
        tickets = [srv.submit(q, b"q%03d" % i) for i, q in enumerate(qs)]
        res = [srv.wait(t) for t in tickets]
        srv.wait(first)
        deferred, refused = srv.admission()
        assert srv.error() == ""
    finally:
        srv.close()
    for r, w in zip(res[:-1], want):
        assert r.status == 0 and np.array_equal(r.scores, w), r
    assert res[-1].status == sdhash_amd.EBUDGET
    assert refused == 1 and deferred > 0, (deferred, refused)

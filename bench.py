"""Benchmark of the sdbf hot path on one MI355X.

Workload (block-mode search): a 1 GiB synthetic disk image of segment-mix
data is hashed in 16 KiB block mode as eight 128 MiB segments (sdhash -b 16
segment naming), 4096 fragments of 4 KiB (half cut from the image, half fresh
random bytes) are stream-digested, and every fragment is compared with every
image segment (4096 x 8 digest pairs, 4096 x 65536 filter pairs).  One step
is that whole pass with all inputs already resident in HBM.

Prints one JSON line.  `value` is the filter-pair comparison rate of the
compare phase; the digest rate (MB/s, 10^6 bytes) of the image phase is in
`digest_MBps`.  Kernel times come from HIP events recorded on the engine's
stream around every launch.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MIB = 1 << 20
IMAGE = 1 << 30
SEGMENT = 128 * MIB
NFRAG = 4096
FRAG = 4096
HBM_PEAK_GBPS = 8000.0
# MI355X: 256 CUs x 4 SIMDs at 2.4 GHz.  VALU: a wave64 instruction per 4
# cycles per SIMD, 2 instructions (AND, popcount-accumulate) per 32-bit word,
# 64 words per filter pair -> 128 lane-instructions per pair.  MFMA: 10.1e15
# dense FP4 ops/s = 5.05e15 MACs/s, 2048 MACs per pair (one per bit pair);
# the int8 form: 5.0e15 dense int8 ops/s.
VALU_PEAK_PAIRS = 256 * 4 * 2.4e9 * 16 / 128
MFMA_PEAK_PAIRS = 10.1e15 / 2 / 2048
MFMA_I8_PEAK_PAIRS = 5.0e15 / 2 / 2048


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(img, frags, seg_names):
    """The CPU oracle on the host cores: block-mode hashing of a 64 MiB slice
    of the image and the compare of 512 fragments against one 128 MiB
    segment (512 x 8192 filter pairs)."""
    import oracle
    threads = oracle.threads_available()
    sample = img[: 64 * MIB]
    t = time.perf_counter()
    oracle.digest_block(sample, 16384, b"sample", threads=threads)
    dt_digest = time.perf_counter() - t
    seg = oracle.digest_block(img[:SEGMENT], 16384, seg_names[0], threads=threads)
    q = oracle.digest_stream_many(frags[:512], [b"frag%04d" % i for i in range(512)], threads=threads)
    q = [d for d in q if d is not None]
    t = time.perf_counter()
    oracle.compare_to(q, [seg], threads=threads)
    dt_cmp = time.perf_counter() - t
    pairs = sum(d.bf_count for d in q) * seg.bf_count
    return {
        "value": pairs / dt_cmp, "unit": "filter-pair cmp/s", "cores": threads, "kind": "port",
        "digest_MBps": sample.size / 1e6 / dt_digest,
        "sample": "oracle (C, OpenMP) on the host: -b 16 hashing of 64 MiB of the image; "
                  "512 fragments x one 128 MiB segment (%d filter pairs)" % pairs,
    }


TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "bench_block_search_1GiB_hbm_traffic.md")


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by
    the copy calibration) from the committed rocprofv3 --pmc summary, or None."""
    try:
        with open(TRAFFIC_TABLE) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) >= 6 and cells[0] == kernel:
                    return float(cells[5])
    except (OSError, ValueError):
        pass
    return None


def _hip():
    import ctypes as C
    h = C.CDLL("libamdhip64.so")
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemcpy.restype = C.c_int
    return h


def download(hip, dbuf, offset: int, nbytes: int) -> np.ndarray:
    """Copies nbytes of a device buffer to the host (hipMemcpyDeviceToHost)."""
    out = np.empty(nbytes, np.uint8)
    if hip.hipMemcpy(out.ctypes.data, dbuf.ptr.value + offset, nbytes, 2) != 0:
        raise RuntimeError("hipMemcpy failed")
    return out


def pct(xs, q):
    return float(np.percentile(np.asarray(xs) * 1e3, q))


def serving(eng, args) -> dict:
    """Serving: a reference set digested on the device from ref_gb of
    device-generated segment mix (1 GiB inputs hashed with -b 10, about one
    filter per 10 KiB), kept resident; queries of 4 KiB-16 MiB (log-uniform),
    half cut from the reference data and half fresh, stream-digested.
    plain: one query at a time through digest + compare_to.  batched: all
    queries submitted to the server at once.  poisson: arrivals at a fixed
    rate (70 % of the batched rate), latency from submit to result."""
    import threading
    import sdhash_amd
    from sdhash_amd import synth
    hip = _hip()
    GIB = 1 << 30
    nfile = max(1, int(round(args.serve_ref_gb * 1e9 / GIB)))
    dref = eng.alloc(nfile * GIB)
    synth.device_fill(dref.ptr.value, nfile * GIB, seed=11)
    ref = eng.new_set()
    t = time.perf_counter()
    st = ref.digest_device(dref, [i * GIB for i in range(nfile)], [GIB] * nfile,
                           [b"ref%04d" % i for i in range(nfile)], dd_block=10240)
    eng.sync()
    t_ref = time.perf_counter() - t
    assert np.all(st == 0)
    text = text_form(eng, ref)
    rng = np.random.Generator(np.random.PCG64(12))
    nq = args.serve_queries
    sizes = np.minimum((2.0 ** rng.uniform(12, 24, nq)).astype(np.int64), 16 * MIB - 16)
    qs = []
    for k, n in enumerate(sizes):
        if k % 2 == 0:                   # cut from the reference data
            o = int(rng.integers(0, nfile * GIB - n)) & ~15
            qs.append(download(hip, dref, o, int(n)))
        else:                            # fresh bytes of another stream
            qs.append(None)
    dref.free()
    fresh_bytes = int(sum(int(n + 15) & ~15 for k, n in enumerate(sizes) if k % 2))
    dq = eng.alloc(max(fresh_bytes, 16))
    synth.device_fill(dq.ptr.value, max(fresh_bytes, 16) + 15 & ~15, seed=13)
    o = 0
    for k, n in enumerate(sizes):
        if k % 2:
            qs[k] = download(hip, dq, o, int(n))
            o += int(n + 15) & ~15
    dq.free()
    names = [b"q%05d" % k for k in range(nq)]

    # plain: one query at a time
    nplain = min(nq, args.serve_plain)
    plain_lat, plain_scores = [], []
    t0 = time.perf_counter()
    for k in range(nplain):
        t = time.perf_counter()
        qset = eng.new_set()
        qset.digest([qs[k]], [names[k]])
        plain_scores.append(eng.compare_to(qset, ref)[0].copy())
        qset.close()
        plain_lat.append(time.perf_counter() - t)
    plain_s = time.perf_counter() - t0

    srv = sdhash_amd.Server(ref, batch_bytes=args.serve_batch_mib * MIB, batch_max=1024)
    # batched: everything at once
    t0 = time.perf_counter()
    tickets = [srv.submit(qs[k], names[k]) for k in range(nq)]
    res = [srv.wait(tk) for tk in tickets]
    batched_s = time.perf_counter() - t0
    nb_closed, _ = srv.stats()
    for k in range(nplain):
        if res[k].status != 0 or not np.array_equal(res[k].scores, plain_scores[k]):
            raise RuntimeError("served result of query %d differs from the plain one" % k)
    # poisson arrivals at fixed fractions of the batched rate
    def poisson(frac):
        rate = frac * nq / batched_s
        gaps = rng.exponential(1.0 / rate, nq)
        tick = [None] * nq

        def submitter():
            tnext = time.perf_counter()
            for k in range(nq):
                tnext += gaps[k]
                d = tnext - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tick[k] = srv.submit(qs[k], names[k])
        th = threading.Thread(target=submitter)
        nb0, _ = srv.stats()
        t0 = time.perf_counter()
        th.start()
        lat = []
        for k in range(nq):
            while tick[k] is None:
                time.sleep(0.0002)
            r = srv.wait(tick[k])
            lat.append(r.latency_s)
            if k < nplain and not np.array_equal(r.scores, plain_scores[k]):
                raise RuntimeError("served result of query %d differs from the plain one" % k)
        th.join()
        el = time.perf_counter() - t0
        nb1, _ = srv.stats()
        return {"offered_qps": rate, "qps": nq / el, "batches": nb1 - nb0,
                "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99)}
    loads = {"40pct": poisson(0.4), "70pct": poisson(0.7)}
    nb_total, _, nb_mfma = srv.stats3()
    srv.close()
    qfilters = sum(int(x.size) for x in qs) / 1e6
    return {
        "ref_filters": ref.filter_count(), "ref_digests": len(ref), "ref_bytes": nfile * GIB,
        "ref_digest_s": t_ref, "queries": nq, "query_MB": qfilters,
        "plain_queries": nplain, "plain_qps": nplain / plain_s,
        "plain_p50_ms": pct(plain_lat, 50), "plain_p99_ms": pct(plain_lat, 99),
        "batched_qps": nq / batched_s, "batched_batches": nb_closed,
        "batched_vs_plain": (nq / batched_s) / (nplain / plain_s),
        "poisson": loads, "mfma_batches": nb_mfma,
        "poisson_target_qps": loads["70pct"]["offered_qps"], "poisson_qps": loads["70pct"]["qps"],
        "p50_ms": loads["70pct"]["p50_ms"], "p99_ms": loads["70pct"]["p99_ms"],
        "identical_results": True,
        "data": "device-generated segment mix (sdbfx_synth_fill); reference -b 10 over 1 GiB inputs",
        "text_form": text,
    }


def text_form(eng, ref) -> dict:
    """The sdbf text of the whole reference set (sdbf_set::to_string), base64
    written by the device, against the oracle's encoder on one of its
    digests on the host; the first digest's text must agree."""
    import oracle
    eng.sync()
    eng.profiling(True)
    eng.reset_kernel_times()
    t = time.perf_counter()
    txt = ref.to_text()
    wall = time.perf_counter() - t
    eng.profiling(False)
    kms = sum(eng.kernel_time(k)[0] for k in ("text_header_kernel", "text_b64_kernel"))
    inf = ref.info(0)
    f, ec, _ = ref.read(0)
    od = oracle.Digest(inf.name.encode("utf-8", "surrogateescape"), f, inf.bf_count, inf.last_count,
                       ec if inf.dd_block_size else None, inf.orig_size, inf.dd_block_size)
    t = time.perf_counter()
    one = od.to_string().encode()
    cpu = time.perf_counter() - t
    if txt[:len(one)] != one:
        raise RuntimeError("device text differs from the oracle's")
    # and read back into a new set (sdbf_set's file reader), bodies decoded
    # by the device; the first and last digests must come back unchanged
    back = eng.new_set()
    eng.sync()
    eng.profiling(True)
    eng.reset_kernel_times()
    t = time.perf_counter()
    n = back.add_text(txt)
    pwall = time.perf_counter() - t
    eng.profiling(False)
    pms = sum(eng.kernel_time(k)[0] for k in ("text_decode_kernel", "text_finish_kernel"))
    for i in (0, len(ref) - 1):
        a, b = ref.read(i), back.read(i)
        if n != len(ref) or any(not np.array_equal(x, y) for x, y in zip(a, b)):
            raise RuntimeError("text read back differs from the set written")
    back.close()
    return {"filters": ref.filter_count(), "text_bytes": len(txt), "kernel_ms": kms,
            "kernel_GBps": len(txt) / (kms / 1e3) / 1e9, "wall_s": wall,
            "cpu_oracle_MBps": len(one) / cpu / 1e6, "cpu_sample": "one digest, %d filters, 1 thread" % inf.bf_count,
            "parse_kernel_ms": pms, "parse_kernel_GBps": len(txt) / (pms / 1e3) / 1e9, "parse_wall_s": pwall}


def _sample_check(eng, s, st, dbuf, offs, sizes, dd, idx, names):
    """Digests idx of the inputs behind set s (statuses st of its
    digest_device call) against the oracle's digests of the same bytes,
    downloaded from dbuf, on every host thread."""
    import oracle
    hip = _hip()
    th = oracle.threads_available()
    pos = np.cumsum(np.asarray(st) == 0) - 1          # set index of each input
    stream = [i for i in idx if dd[i] == 0]
    bufs = [download(hip, dbuf, int(offs[i]), int(sizes[i])) for i in stream]
    ods = dict(zip(stream, oracle.digest_stream_many(bufs, [names[i] for i in stream], threads=th)))
    for i in idx:
        if dd[i]:
            ods[i] = oracle.digest_block(download(hip, dbuf, int(offs[i]), int(sizes[i])), int(dd[i]), names[i],
                                         threads=th)
    for i in idx:
        od = ods[i]
        if od is None:
            if st[i] != -3:
                raise RuntimeError("input %d: the oracle refuses it, the engine does not" % i)
            continue
        f, ec, _ = s.read(int(pos[i]))
        inf = s.info(int(pos[i]))
        same = st[i] == 0 and np.array_equal(f, od.filters) and inf.orig_size == od.orig_size
        if od.elem_counts is not None:
            same = same and np.array_equal(ec, od.elem_counts)
        else:
            same = same and inf.last_count == od.last_count
        if not same:
            raise RuntimeError("digest of input %d differs from the oracle's" % i)
    return len(idx), int(sum(int(sizes[i]) for i in idx))


def batched_digest(eng, args) -> dict:
    """Batched digesting: `batch_files` files (76,000 by default, about 60 GB)
    with log-normal sizes (median 256 KiB, sigma 1.5, clamped to 4 KiB-64
    MiB) of device-generated segment mix, 5 % of each file's bytes planted as
    4-64 KiB chunks copied from a shared 1 GiB donor region (so that files
    share content), resident in HBM; >= 16 MiB hashed with -b 16, the rest in
    stream mode (the sdhash CLI policy).  batched: one call for all files;
    plain: one call per file over the first `plain` files.  Checked against
    the oracle on a sample of files of every size, and batched against
    per-file on a sample."""
    from sdhash_amd import synth
    rng = np.random.Generator(np.random.PCG64(21))
    n = args.batch_files
    sizes = np.clip(np.exp(rng.normal(np.log(256 * 1024), 1.5, n)), 4096, 64 * MIB).astype(np.int64)
    offs = np.zeros(n, np.int64)
    offs[1:] = np.cumsum((sizes[:-1] + 255) // 256 * 256)
    total = int(offs[-1] + (sizes[-1] + 255) // 256 * 256)
    dd = np.where(sizes >= 16 * MIB, 16384, 0).astype(np.uint32)
    dbuf = eng.alloc(total)
    synth.device_fill(dbuf.ptr.value, total, seed=31)
    # planted chunks: non-overlapping slots of each file, from the donor
    DONOR = 1 << 30
    donor = eng.alloc(DONOR)
    synth.device_fill(donor.ptr.value, DONOR, seed=32)
    d_off, s_off, lens = [], [], []
    for i in range(n):
        want = int(0.05 * sizes[i])
        if want < 4096:
            continue
        c = rng.integers(4096, 65536 + 1, max(1, want // 4096)) & ~15
        c = c[np.cumsum(c) <= want] if (np.cumsum(c) <= want).any() else c[:1]
        slot = int(sizes[i]) // len(c)
        for k, ck in enumerate(c):
            ck = int(min(ck, slot)) & ~15
            if ck < 16:
                continue
            d_off.append(int(offs[i]) + ((k * slot) & ~15))
            s_off.append(int(rng.integers(0, DONOR - ck)) & ~15)
            lens.append(ck)
    synth.device_plant(dbuf.ptr.value, total, donor.ptr.value, DONOR, d_off, s_off, lens)
    donor.free()
    names = [b"file%05d" % i for i in range(n)]
    s = eng.new_set()
    eng.sync()
    t = time.perf_counter()
    st = s.digest_device(dbuf, offs, sizes, names, dd_block=dd)
    eng.sync()
    t_batched = time.perf_counter() - t
    nplain = min(n, args.batch_plain)
    plain = []
    t = time.perf_counter()
    for i in range(nplain):
        p = eng.new_set()
        p.digest_device(dbuf, offs[i:i + 1], sizes[i:i + 1], names[i:i + 1], dd_block=dd[i:i + 1])
        plain.append(p)
    eng.sync()
    t_plain = time.perf_counter() - t
    pos = np.cumsum(st == 0) - 1
    for i in range(0, nplain, max(1, nplain // 200)):
        if st[i] != 0:
            continue
        a, b = s.read(int(pos[i])), plain[i].read(0)
        if not all(np.array_equal(x, y) for x, y in zip(a, b)):
            raise RuntimeError("batched digest of file %d differs from the per-file one" % i)
    plain_bytes = int(sizes[:nplain].sum())
    for p in plain:
        p.close()
    # oracle sample: files spread over the size range, and the largest ones
    order = np.argsort(sizes, kind="stable")
    idx = sorted(set(order[np.linspace(0, n - 1, 120).astype(int)].tolist() + order[-3:].tolist()))
    nsample, sample_bytes = _sample_check(eng, s, st, dbuf, offs, sizes, dd, idx, names)
    dbuf.free()
    return {
        "files": n, "bytes": int(sizes.sum()), "block_mode_files": int((dd > 0).sum()),
        "planted_chunks": len(lens), "planted_bytes": int(sum(lens)),
        "batched_s": t_batched, "batched_GBps": float(sizes.sum()) / t_batched / 1e9,
        "plain_files": nplain, "plain_s": t_plain, "plain_GBps": plain_bytes / t_plain / 1e9,
        "batched_vs_plain": (float(sizes.sum()) / t_batched) / (plain_bytes / t_plain),
        "too_small": int((st == -3).sum()), "identical_sample": True,
        "oracle_sample": {"files": nsample, "bytes": sample_bytes, "identical": True},
        "data": "device-generated segment mix (sdbfx_synth_fill) with 5 % planted donor chunks",
    }


def allpairs(eng, args) -> dict:
    """All-pairs compare: about `allpairs_mfilters` million filters (2 by
    default) in digests of 1-6,000 filters (uniform), hashed with -b 16 from
    segments cut at 16 KiB-aligned random offsets of one device-generated
    image half their total size, so that overlapping segments share blocks
    and filters; compare_all with each comparison kernel, same scores
    required.  Rates are filter pairs per second of kernel time.  Checked
    against the oracle: a sample of digests digested on the host, and the
    scores of a sample of digest pairs."""
    import oracle
    import sdhash_amd
    from sdhash_amd import synth
    rng = np.random.Generator(np.random.PCG64(41))
    target = int(args.allpairs_mfilters * 1e6)
    counts = []
    while sum(counts) < target:
        counts.append(int(rng.integers(1, 6001)))
    sizes = np.array(counts, np.int64) * 16384
    img_blocks = max(int(sizes.sum()) // 16384 // 2, max(counts))
    offs = np.array([int(rng.integers(0, img_blocks - c + 1)) * 16384 for c in counts], np.int64)
    dimg = eng.alloc(img_blocks * 16384)
    synth.device_fill(dimg.ptr.value, img_blocks * 16384, seed=43)
    names = [b"d%04d" % i for i in range(len(sizes))]
    s = eng.new_set()
    s.digest_device(dimg, offs, sizes, names, dd_block=16384)
    eng.sync()
    n = len(s)
    nf = np.array([s.info(i).bf_count for i in range(n)], np.float64)
    pairs = float((nf.sum() ** 2 - (nf ** 2).sum()) / 2)
    out = {"digests": n, "filters": s.filter_count(), "filter_pairs": pairs,
           "image_bytes": img_blocks * 16384, "segment_bytes": int(sizes.sum())}
    ref = None
    for label, method, kname in (("mfma", sdhash_amd.CMP_MFMA, "cmp_mfma_kernel"),
                                 ("mfma_i8", sdhash_amd.CMP_MFMA_I8, "cmp_mfma_i8_kernel"),
                                 ("valu", sdhash_amd.CMP_VALU, "cmp_kernel")):
        eng.compare_method(method)
        eng.sync()
        eng.profiling(True)
        eng.reset_kernel_times()
        t = time.perf_counter()
        sc = eng.compare_all(s)
        wall = time.perf_counter() - t
        eng.profiling(False)
        kms = sum(eng.kernel_time(k)[0] for k in (kname, "col_order_kernel", "cmp_prep_kernel", "chunk_max_kernel",
                                                   "pair_sum_kernel"))
        if ref is None:
            ref = sc
        elif not np.array_equal(sc, ref):
            raise RuntimeError("all-pairs scores differ between the comparison kernels")
        out[label] = {"kernel_ms": kms, "wall_s": wall, "pairs_per_s": pairs / (kms / 1e3),
                      "cmp_kernel_ms": eng.kernel_time(kname)[0]}
    eng.compare_method(sdhash_amd.CMP_AUTO)
    out["scores_ge_21"] = int((ref >= 21).sum())
    # oracle: digests of a sample hashed on the host; the scores of all pairs
    # among 12 digests of at most 1,500 filters, and of 8 pairs of digests of
    # at most 3,000 filters that score >= 21 on the device
    _sample_check(eng, s, np.zeros(n, np.int32), dimg, offs, sizes, [16384] * n,
                  sorted(rng.choice(n, min(6, n), replace=False).tolist()), names)
    dimg.free()
    from concurrent.futures import ThreadPoolExecutor
    cnt = np.array(counts)
    ia, ib = np.triu_indices(n, 1)
    hot = np.flatnonzero((ref >= 21) & (cnt[ia] <= 3000) & (cnt[ib] <= 3000))
    hot = rng.permutation(hot)[:8]
    small = rng.permutation(np.flatnonzero(cnt <= 1500))[:12].tolist()
    pl = [(int(ia[k]), int(ib[k])) for k in hot] + [(a, b) for x, a in enumerate(sorted(small))
                                                      for b in sorted(small)[x + 1:]]
    ods = {}
    for i in {x for p in pl for x in p}:
        f, ec, _ = s.read(i)
        inf = s.info(i)
        ods[i] = oracle.Digest(names[i], f, inf.bf_count, 0, ec, inf.orig_size, 16384)
        ods[i].cstruct()
    with ThreadPoolExecutor(oracle.threads_available()) as ex:
        want = list(ex.map(lambda p: oracle.score(ods[p[0]], ods[p[1]]), pl))
    for (a, b), w in zip(pl, want):
        if int(ref[a * n - a * (a + 1) // 2 + (b - a - 1)]) != w:
            raise RuntimeError("score of digests %d, %d differs from the oracle's" % (a, b))
    out["oracle_sample"] = {"digests_hashed": min(6, n), "pairs_scored": len(pl),
                            "pairs_ge_21": int(sum(w >= 21 for w in want)), "identical": True}
    out["data"] = ("device-generated segment mix (sdbfx_synth_fill), -b 16 segments of one image at "
                   "16 KiB-aligned random offsets (overlapping)")
    return out


def index_section(eng, dimg, img) -> dict:
    """sdhash --index / --index-search on the device, over the 1 GiB image in
    HBM: (1) block-mode insertion of its eight 128 MiB segments at -b 16 into
    a 64 MiB index (the --index hashing of large files); (2) the stream-mode
    gate over 64 files of 1 MiB cut at random offsets of the image (so that
    later files repeat features of earlier ones), in order, into a fresh
    index; (3) the image at -b 16 searched against four 64 MiB indexes.  Rates
    are input bytes per second of wall time around each call (inputs resident,
    index scratch allocated by a warm-up call).  Each is checked against the
    oracle: (1) and (3) on a 64 MiB slice of the image, (2) on all 64 files;
    the oracle's times on those are the CPU baselines (one thread: the
    reference's index inserts are ordered)."""
    import oracle
    from sdhash_amd import synth
    nseg = IMAGE // SEGMENT
    seg_off = [i * SEGMENT for i in range(nseg)]
    names = [b"image.%04dM" % (i * SEGMENT // MIB) for i in range(nseg)]
    zeros = np.zeros(64 << 20, np.uint8)
    out = {}

    def timed(fn, reps=3):
        fn()
        eng.sync()
        best = 1e9
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            eng.sync()
            best = min(best, time.perf_counter() - t)
        return best

    # (1) block-mode insertion
    ix = eng.new_index(64 << 20)

    def blk():
        ix.write(zeros, 0)
        s = eng.new_set()
        s.digest_device_indexed(dimg, seg_off, [SEGMENT] * nseg, names, dd_block=16384, index=ix)
        s.close()
    t = timed(blk)
    t_write = timed(lambda: ix.write(zeros, 0))
    blk_s = t - t_write
    # parity and CPU baseline on a 64 MiB slice
    sl = 64 * MIB
    ix.write(zeros, 0)
    s = eng.new_set()
    s.digest_device_indexed(dimg, [0], [sl], [b"slice"], dd_block=16384, index=ix)
    gbits, gec = ix.read()
    oix = oracle.Index(64 << 20)
    tc = time.perf_counter()
    od = oracle.digest_block_indexed(img[:sl], 16384, b"slice", index=oix)[0]
    cpu_blk = time.perf_counter() - tc
    if gec != oix.elem_count or not np.array_equal(gbits, oix.bf) or not np.array_equal(s.read(0)[0], od.filters):
        raise RuntimeError("indexed block digest differs from the oracle's")
    s.close()
    out["block_insert"] = {"bytes": IMAGE, "wall_s": blk_s, "GBps": IMAGE / blk_s / 1e9,
                           "cpu_oracle_MBps": sl / cpu_blk / 1e6, "cpu_sample": "64 MiB slice, 1 thread",
                           "identical_sample": True}
    # (2) the stream gate, 64 files of 1 MiB
    rng = np.random.Generator(np.random.PCG64(61))
    offs = [int(rng.integers(0, (IMAGE - MIB) // 256)) * 256 for _ in range(64)]
    fnames = [b"f%02d" % k for k in range(64)]
    six = eng.new_index(64 << 20)

    def strm():
        six.write(zeros, 0)
        s = eng.new_set()
        s.digest_device_indexed(dimg, offs, [MIB] * 64, fnames, dd_block=0, index=six)
        s.close()
    t = timed(strm)
    t_write = timed(lambda: six.write(zeros, 0))
    st_s = t - t_write
    six.write(zeros, 0)
    s = eng.new_set()
    s.digest_device_indexed(dimg, offs, [MIB] * 64, fnames, dd_block=0, index=six)
    gbits, gec = six.read()
    oix = oracle.Index(64 << 20)
    tc = time.perf_counter()
    ods = [oracle.digest_stream_indexed(img[o:o + MIB], oix, n) for o, n in zip(offs, fnames)]
    cpu_st = time.perf_counter() - tc
    same = gec == oix.elem_count and np.array_equal(gbits, oix.bf)
    for k, od in enumerate(ods):
        same = same and s.info(k).bf_count == od.bf_count and np.array_equal(s.read(k)[0], od.filters)
    if not same:
        raise RuntimeError("indexed stream digests differ from the oracle's")
    s.close()
    out["stream_gate"] = {"files": 64, "bytes": 64 * MIB, "wall_s": st_s, "MBps": 64 * MIB / st_s / 1e6,
                          "cpu_oracle_MBps": 64 * MIB / cpu_st / 1e6, "cpu_sample": "all 64 files, 1 thread",
                          "identical": True}
    # (3) search against four indexes: the image's own (1) and three built
    # from 256 MiB regions of fresh device data
    ixs = [ix]
    gen = eng.alloc(256 * MIB)
    for k in range(3):
        synth.device_fill(gen.ptr.value, 256 * MIB, seed=71 + k)
        x = eng.new_index(64 << 20)
        s = eng.new_set()
        s.digest_device_indexed(gen, [0, 128 * MIB], [128 * MIB] * 2, dd_block=16384, index=x)
        s.close()
        ixs.append(x)
    gen.free()
    ix.write(zeros, 0)
    s = eng.new_set()
    s.digest_device_indexed(dimg, seg_off, [SEGMENT] * nseg, names, dd_block=16384, index=ix)
    s.close()

    def srch():
        s = eng.new_set()
        st, m = s.digest_device_indexed(dimg, seg_off, [SEGMENT] * nseg, names, dd_block=16384, search=ixs)
        s.close()
        return m
    sr_s = timed(srch)
    m = srch()
    host = [x.read() for x in ixs]
    hix = [oracle.Index(64 << 20, b, e) for b, e in host]
    tc = time.perf_counter()
    _, om = oracle.digest_block_indexed(img[:sl], 16384, b"q", search=hix)
    cpu_sr = time.perf_counter() - tc
    if not np.array_equal(m[:om.shape[0]], om):
        raise RuntimeError("index search counts differ from the oracle's")
    for x in ixs:
        x.close()
    six.close()
    out["search"] = {"bytes": IMAGE, "indexes": 4, "wall_s": sr_s, "GBps": IMAGE / sr_s / 1e9,
                     "blocks_over_4": int((m > 4).sum()), "cpu_oracle_MBps": sl / cpu_sr / 1e6,
                     "cpu_sample": "64 MiB slice against the four indexes, 1 thread", "identical_sample": True}
    return out


def graph_replay(eng, dimg) -> dict:
    """The no-allocation block digest of a 1 MiB window (64 dd-blocks of
    16 KiB) called 500 times back to back: eager calls
    (sdbfx_digest_blocks_noalloc, two launches each) against replays of the
    same launches captured once into a HIP graph (sdbfx_graph_launch).  Rates
    are calls per second of wall time, synchronised once at the end; both
    outputs are compared."""
    import ctypes as C
    hip = _hip()
    win, dd, calls = MIB, 16384, 500
    n = win // dd
    outs = [(eng.alloc(n * 256), eng.alloc(n * 2), eng.alloc(n * 2)) for _ in range(2)]
    nf = C.c_uint32()
    g = C.c_void_p()
    fb, eb, hb = outs[1]
    if eng.lib.sdbfx_graph_blocks_create(eng.ctx, dimg.ptr, win, dd, fb.ptr, eb.ptr, hb.ptr, C.byref(nf),
                                         C.byref(g)) != 0:
        raise RuntimeError(eng.lib.sdbfx_last_error(eng.ctx).decode())
    fa, ea, ha = outs[0]

    def eager():
        for _ in range(calls):
            eng.lib.sdbfx_digest_blocks_noalloc(eng.ctx, dimg.ptr, win, dd, fa.ptr, ea.ptr, ha.ptr, C.byref(nf))
        eng.sync()

    def graph():
        for _ in range(calls):
            eng.lib.sdbfx_graph_launch(eng.ctx, g)
        eng.sync()
    res = {}
    for name, fn in (("eager", eager), ("graph", graph)):
        fn()
        t = time.perf_counter()
        fn()
        res[name] = calls / (time.perf_counter() - t)
    same = all(np.array_equal(download(hip, a, 0, n * w), download(hip, b, 0, n * w))
               for (a, b, w) in zip(outs[0], outs[1], (256, 2, 2)))
    eng.lib.sdbfx_graph_destroy(g)
    for o in outs:
        for b in o:
            b.free()
    if not same:
        raise RuntimeError("graph replay output differs from the eager call's")
    return {"window_bytes": win, "calls": calls, "eager_calls_per_s": res["eager"], "graph_calls_per_s": res["graph"],
            "graph_vs_eager": res["graph"] / res["eager"], "identical": True}


def plain_allpairs(eng) -> dict:
    """The plain method: 64 files of 1 MiB of segment mix with 0-50 %
    planted cross-file chunks (host-generated), uploaded to HBM, stream
    digests and all 2,016 pair scores (sdhash -g); the oracle on every host
    thread does the same and must agree."""
    import oracle
    from sdhash_amd import synth
    files = synth.file_set(101, 64, MIB)
    names = [b"dir/file%02d.bin" % i for i in range(64)]
    dbuf = eng.alloc(64 * MIB)
    dbuf.upload(0, np.concatenate(files))
    offs = [i * MIB for i in range(64)]
    steps = 5
    eng.sync()
    eng.profiling(True)
    eng.reset_kernel_times()
    t = time.perf_counter()
    for _ in range(steps):
        s = eng.new_set()
        s.digest_device(dbuf, offs, [MIB] * 64, names, dd_block=0)
        sc = eng.compare_all(s)
    eng.sync()
    wall = (time.perf_counter() - t) / steps
    eng.profiling(False)
    dig_ms = sum(eng.kernel_time(k)[0] for k in ("stream_tile_kernel", "stream_fixup_plan_kernel",
                                                 "stream_fixup_par_kernel", "stream_fixup_kernel",
                                                 "stream_insert_kernel", "gather_filters_kernel")) / steps
    cmp_ms = sum(eng.kernel_time(k)[0] for k in ("cmp_kernel", "cmp_mfma_kernel", "col_order_kernel",
                                                 "cmp_prep_kernel", "chunk_max_kernel", "pair_sum_kernel")) / steps
    nf = np.array([s.info(i).bf_count for i in range(64)], np.float64)
    pairs = float((nf.sum() ** 2 - (nf ** 2).sum()) / 2)
    dbuf.free()
    th = oracle.threads_available()
    t = time.perf_counter()
    ods = oracle.digest_stream_many(files, names, threads=th)
    cpu_dig = time.perf_counter() - t
    t = time.perf_counter()
    want = oracle.compare_all(ods, threads=th)
    cpu_cmp = time.perf_counter() - t
    if not np.array_equal(sc, want):
        raise RuntimeError("plain all-pairs scores differ from the oracle's")
    for i in range(64):
        if not np.array_equal(s.read(i)[0], ods[i].filters):
            raise RuntimeError("plain digest %d differs from the oracle's" % i)
    return {"files": 64, "bytes": 64 * MIB, "filters": int(nf.sum()), "filter_pairs": pairs, "scores": int(sc.size),
            "step_wall_ms": wall * 1e3, "digest_kernel_ms": dig_ms, "compare_kernel_ms": cmp_ms,
            "digest_MBps": 64 * MIB / 1e6 / (dig_ms / 1e3), "compare_pairs_per_s": pairs / (cmp_ms / 1e3),
            "cpu_baseline": {"digest_MBps": 64 * MIB / 1e6 / cpu_dig, "compare_pairs_per_s": pairs / cpu_cmp,
                             "cores": th, "kind": "port", "sample": "the whole workload"},
            "identical": True, "scores_ge_1": int((want >= 1).sum())}


def copy_probe() -> int:
    """A plain 256 MiB device-to-device copy, three times: the reference
    point for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.ones(256 * MIB, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied 3 x %d bytes" % a.numel())
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-digest measurement")
    ap.add_argument("--serve-ref-gb", type=float, default=100.0)
    ap.add_argument("--serve-queries", type=int, default=256)
    ap.add_argument("--serve-plain", type=int, default=256)
    ap.add_argument("--serve-batch-mib", type=int, default=512)
    ap.add_argument("--batch-files", type=int, default=76000)
    ap.add_argument("--batch-plain", type=int, default=4000)
    ap.add_argument("--no-allpairs", action="store_true", help="skip the all-pairs measurement")
    ap.add_argument("--allpairs-mfilters", type=float, default=2.0)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain 64 x 1 MiB all-pairs measurement")
    ap.add_argument("--no-index", action="store_true", help="skip the --index / --index-search measurement")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP graph replay measurement")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 256 MiB device copy (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        return 2

    import sdhash_amd
    from sdhash_amd import synth

    t0 = time.perf_counter()
    img = synth.disk_image(1, IMAGE)
    frags = synth.fragments(2, img, NFRAG, FRAG)
    log("inputs generated in %.1f s" % (time.perf_counter() - t0))

    eng = sdhash_amd.Engine(0)
    # the comparison configuration of the step is chosen by the autotuner
    # (the first warm-up step times every candidate and keeps the fastest)
    eng.autotune(True)
    dimg = eng.alloc(IMAGE)
    dimg.upload(0, img)
    fpitch = (FRAG + 255) // 256 * 256
    dfr = eng.alloc(NFRAG * fpitch)
    dfr.upload(0, np.stack(frags).reshape(NFRAG, FRAG) if fpitch == FRAG else
               np.concatenate([np.pad(f, (0, fpitch - FRAG)) for f in frags]))
    nseg = IMAGE // SEGMENT
    seg_off = [i * SEGMENT for i in range(nseg)]
    seg_len = [SEGMENT] * nseg
    seg_names = [b"image.%04dM" % (i * SEGMENT // MIB) for i in range(nseg)]
    # the fragments' offsets, lengths and names as the arrays the library
    # takes, built once rather than converted from Python lists every step
    fr_off = np.arange(NFRAG, dtype=np.uint64) * fpitch
    fr_len = np.full(NFRAG, FRAG, np.uint64)
    fr_names = sdhash_amd.name_array([b"frag%04d" % i for i in range(NFRAG)])

    def step():
        t = eng.new_set()
        st = t.digest_device(dimg, seg_off, seg_len, seg_names, dd_block=16384)
        q = eng.new_set()
        sq = q.digest_device(dfr, fr_off, fr_len, fr_names, dd_block=0)
        scores = eng.compare_to(q, t)
        return t, q, scores, st, sq

    for _ in range(args.warmup):
        step()
    eng.sync()
    eng.profiling(True)
    eng.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tset, qset, scores, st, sq = step()
    eng.sync()
    elapsed = time.perf_counter() - t0
    eng.profiling(False)

    k = {}
    for name in ("block_digest_kernel", "stream_tile_kernel", "stream_fixup_plan_kernel", "stream_fixup_par_kernel",
                 "stream_fixup_kernel", "stream_insert_kernel", "gather_filters_kernel", "cmp_kernel",
                 "col_order_kernel", "cmp_prep_kernel", "cmp_mfma_kernel", "chunk_max_kernel", "pair_sum_kernel"):
        ms, n = eng.kernel_time(name)
        if n:
            k[name] = (ms, n)
    nfilt_t = tset.filter_count()
    pairs = sum(qset.info(i).bf_count for i in range(len(qset))) * nfilt_t
    cmp_ms = sum(k.get(n, (0.0, 0))[0] for n in ("cmp_kernel", "col_order_kernel", "cmp_prep_kernel",
                                                 "cmp_mfma_kernel", "chunk_max_kernel", "pair_sum_kernel")) / args.steps

    # the comparison kernels on the same sets: same scores, rate of each
    # against the peak pair rate of its unit (2048 AND-popcount bit pairs per
    # filter pair: VALU 2 instructions per 32-bit word per lane; MFMA 2048
    # FP4 MACs, or 2048 int8 MACs in the byte-plane form)
    forms = {}
    for label, method, kname, peak in (
            ("valu", sdhash_amd.CMP_VALU, "cmp_kernel", VALU_PEAK_PAIRS),
            ("mfma", sdhash_amd.CMP_MFMA, "cmp_mfma_kernel", MFMA_PEAK_PAIRS),
            ("mfma_i8", sdhash_amd.CMP_MFMA_I8, "cmp_mfma_i8_kernel", MFMA_I8_PEAK_PAIRS)):
        eng.compare_method(method)
        got = eng.compare_to(qset, tset)
        if not np.array_equal(got, scores):
            log("compare method %s disagrees with the step's scores" % label)
            return 1
        eng.sync()
        eng.profiling(True)
        eng.reset_kernel_times()
        for _ in range(args.steps):
            eng.compare_to(qset, tset)
        eng.sync()
        eng.profiling(False)
        kms, kn = eng.kernel_time(kname)
        sms, _ = eng.kernel_time("pair_sum_kernel")
        rate = pairs / (kms / max(kn, 1) / 1e3)
        forms[label] = {"kernel": kname, "kernel_ms": kms / max(kn, 1), "pair_sum_ms": sms / max(kn, 1),
                        "pairs_per_s": rate, "peak_pairs_per_s": peak, "frac_of_unit_peak": rate / peak}
    eng.compare_method(sdhash_amd.CMP_AUTO)
    # fast mode (sdhash --fast): both sets folded to 64-byte filters, scored
    # by the 16-word form of the VALU kernel (a quarter of the AND-popcounts)
    qset.fold()
    tset.fold()
    eng.compare_to(qset, tset)
    eng.sync()
    eng.profiling(True)
    eng.reset_kernel_times()
    for _ in range(args.steps):
        eng.compare_to(qset, tset)
    eng.sync()
    eng.profiling(False)
    kms, kn = eng.kernel_time("cmp_fast_kernel")
    forms["fast"] = {"kernel": "cmp_fast_kernel", "kernel_ms": kms / max(kn, 1),
                     "pairs_per_s": pairs / (kms / max(kn, 1) / 1e3),
                     "note": "64-byte folded filters: not the same scores as the 256-byte forms"}
    # the MFMA kernel's rate in the step (with the autotuned configuration),
    # or the forced form's when the step did not use it
    mfma_rate = (pairs / (k["cmp_mfma_kernel"][0] / k["cmp_mfma_kernel"][1] / 1e3) if "cmp_mfma_kernel" in k
                 else forms["mfma"]["pairs_per_s"])
    blk_ms, blk_n = k["block_digest_kernel"]
    blk_avg = blk_ms / max(blk_n, 1)
    digest_ms = blk_ms / args.steps
    ms_step = elapsed * 1000 / args.steps
    achieved = (IMAGE * args.steps / max(blk_n, 1)) / (blk_avg / 1e3) / 1e9

    out = {
        "metric": "filter-pair cmp/s + MB/s digested, bit-exact; batched/served vs plain; p99 ms",
        "value": pairs / (cmp_ms / 1e3),
        "unit": "filter-pair cmp/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8/u32 integer (f64 final score)", "data": "synthetic segment mix",
        "config": {"workload": "1 GiB image -b 16 (8 x 128 MiB segments) + 4096 x 4 KiB fragment search",
                   "filter_pairs_per_step": pairs, "image_filters": nfilt_t},
        "digest_MBps": IMAGE / 1e6 / (digest_ms / 1e3),
        "kernel_ms_per_step": {n: v[0] / args.steps for n, v in k.items()},
        "compare_forms": forms,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": measured_traffic("block_digest_kernel"),
                     "kernel": "block_digest_kernel",
                     "algorithmic_bytes_per_launch": IMAGE * args.steps // max(blk_n, 1)},
        # the kernel behind `value`: 2048 FP4 MACs (4096 ops) per filter pair
        "roofline_compare": {"bound": "mfma", "kernel": "cmp_mfma_kernel", "unit": "TFLOP/s",
                             "dtype": "fp4",
                             "achieved": mfma_rate * 4096 / 1e12, "peak": 10100.0,
                             "frac": mfma_rate * 4096 / 1e12 / 10100.0,
                             "traffic": measured_traffic("cmp_mfma_kernel"),
                             "algorithmic_ops_per_launch": pairs * 4096,
                             "launch_ms": pairs / mfma_rate * 1e3},
        "autotune": eng.tuned(),
    }
    del tset, qset
    # the other sections use the engine's cost model, as the server does
    eng.autotune(False)
    if not args.no_graph:
        out["graph_replay"] = graph_replay(eng, dimg)
    if not args.no_index:
        log("index ...")
        out["index"] = index_section(eng, dimg, img)
    if not args.no_plain:
        log("plain all-pairs ...")
        out["plain_allpairs"] = plain_allpairs(eng)
    if not args.no_batched:
        log("batched digest ...")
        out["batched_digest"] = batched_digest(eng, args)
    if not args.no_allpairs:
        log("all-pairs ...")
        out["allpairs"] = allpairs(eng, args)
    if not args.no_serving:
        log("serving ...")
        out["serving"] = serving(eng, args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(img, frags, seg_names)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

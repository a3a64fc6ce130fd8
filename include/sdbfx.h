/*
 * sdbfx.h -- C ABI of the MI355X (gfx950) sdbf engine.
 *
 * Plain pointers and sizes only.  Every call returns 0 on success or a
 * negative status; sdbfx_last_error() describes the last failure of a
 * context.  A context owns one HIP stream on one device and is not
 * thread-safe: use one context per host thread.
 *
 * Status codes follow the reference's exception codes
 * (sdbf/sdbf_class.cc:37,81,117,146,153,192,209):
 *    0  ok
 *   -1  cannot process (bad arguments, device failure, out of memory)
 *   -2  invalid digest format
 *   -3  input too small (< 512 bytes)
 *   -4  the call's scratch would exceed the context's memory budget
 *       (SDBFX_EBUDGET); nothing was done
 *
 * The reference binds this path through C++ classes (sdbf, sdbf_set); each
 * entry point below names the reference function it replaces.
 */
#ifndef SDBFX_H
#define SDBFX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sdbfx_ctx sdbfx_ctx;   /* device + stream + scratch buffers */
typedef struct sdbfx_set sdbfx_set;   /* device-resident collection of digests */
typedef struct sdbfx_index sdbfx_index; /* device-resident index (sdhash --index) */

/* Per-digest metadata, as held by class sdbf (sdbf/sdbf_class.h:95-117). */
typedef struct {
    const char *name;        /* hashname */
    uint64_t orig_size;      /* orig_file_size */
    uint32_t bf_count;       /* number of 256-byte filters */
    uint32_t last_count;     /* stream mode: elements in the last filter */
    uint32_t dd_block_size;  /* 0 for stream mode, else bytes per dd-block */
    uint32_t max_elem;       /* 160 stream, 192 block */
} sdbfx_digest_info;

/* ---- context ------------------------------------------------------- */

#define SDBFX_EBUDGET (-4)

/* Opens device `device`; mem_budget (bytes, 0 = no limit) bounds the
 * scratch device memory the context holds (staging, work lists, partial
 * results; not the digest sets, which the caller owns): a call that would
 * grow it past the budget returns SDBFX_EBUDGET and changes nothing.
 * Scratch is kept between calls (sdbfx_ctx_trim frees it).  Replaces the
 * process-global sdbf::config set-up (sdbf/sdbf_class.cc:25,
 * sdhash-src/sdhash.cc:227); the reference's only bound is exit(-1) when an
 * allocation fails (sdbf/error.cc:33-38). */
int sdbfx_ctx_create(int device, uint64_t mem_budget, sdbfx_ctx **out);
void sdbfx_ctx_destroy(sdbfx_ctx *ctx);
const char *sdbfx_last_error(const sdbfx_ctx *ctx);
/* waits for all work of the context */
int sdbfx_sync(sdbfx_ctx *ctx);
/* scratch bytes the context holds now */
uint64_t sdbfx_ctx_scratch_bytes(const sdbfx_ctx *ctx);
/* frees all scratch of the context (after waiting for its work) */
int sdbfx_ctx_trim(sdbfx_ctx *ctx);

/* ---- digest sets ---------------------------------------------------- */

/* Empty set (sdbf_set::sdbf_set(), sdbf/sdbf_set.cc:29-34). */
int sdbfx_set_create(sdbfx_ctx *ctx, sdbfx_set **out);
void sdbfx_set_destroy(sdbfx_set *set);
uint32_t sdbfx_set_size(const sdbfx_set *set);
uint64_t sdbfx_set_filter_count(const sdbfx_set *set);

/*
 * Digests n host buffers and appends the digests, in order, to `set`.
 * dd_block[i] == 0 selects stream mode, else block mode with that dd-block
 * size in bytes (a multiple of 16 in [512, 16384]; sdhash's -b N gives
 * N KiB); dd_block == NULL means stream mode for all.
 * status[i] receives 0 or -3 (too small: nothing appended for that input).
 * Replaces sdbf::sdbf(name, char*, dd, len, NULL) (sdbf/sdbf_class.cc:115-136)
 * and sdbf::sdbf(name, istream*, dd, size, NULL) (sdbf/sdbf_class.cc:71-102)
 * as called in batches by sdbf_hash_files / sdbf_hash_files_dd
 * (sdhash-src/sdhash_threads.cc:241-373) followed by sdbf_set::add
 * (sdbf/sdbf_set.cc:105-109).
 */
int sdbfx_set_digest(sdbfx_ctx *ctx, sdbfx_set *set, uint32_t n, const uint8_t *const *bufs,
                     const uint64_t *lens, const char *const *names, const uint32_t *dd_block,
                     int32_t *status);

/*
 * As sdbfx_set_digest, for inputs already in device memory: input i is at
 * d_base + offsets[i].  Offsets must be multiples of 256 and the 256 bytes
 * after each input must be readable (sdbfx_device_alloc pads for this).
 */
int sdbfx_set_digest_device(sdbfx_ctx *ctx, sdbfx_set *set, uint32_t n, const void *d_base,
                            const uint64_t *offsets, const uint64_t *lens, const char *const *names,
                            const uint32_t *dd_block, int32_t *status);

/*
 * Appends one already-computed digest (e.g. parsed from sdbf text).
 * elem_counts is required for block mode (dd_block_size > 0) and ignored
 * for stream mode.  Replaces sdbf::sdbf(const std::string&)'s fill of the
 * object (sdbf/sdbf_class.cc:144-215) followed by sdbf_set::add.
 */
int sdbfx_set_add(sdbfx_ctx *ctx, sdbfx_set *set, const char *name, uint64_t orig_size,
                  uint32_t bf_count, uint32_t last_count, uint32_t dd_block_size, uint32_t max_elem,
                  const uint8_t *filters, const uint16_t *elem_counts);

/* Metadata of digest idx (sdbf::name/size/input_size/filter_count,
 * sdbf/sdbf_class.cc:236-256,418-420).  info->name stays valid while the set lives. */
int sdbfx_set_info(const sdbfx_set *set, uint32_t idx, sdbfx_digest_info *info);

/* Copies digest idx back to the host: bf_count*256 filter bytes, bf_count
 * element counts (sdbf::get_elem_count, sdbf/sdbf_class.cc:407-415) and
 * bf_count Hamming weights (sdbf::compute_hamming, sdbf/sdbf_class.cc:389-401).
 * Any output pointer may be NULL. */
int sdbfx_set_read(sdbfx_ctx *ctx, const sdbfx_set *set, uint32_t idx, uint8_t *filters,
                   uint16_t *elem_counts, uint16_t *hamming);

/*
 * The text form of every digest of the set, in set order, exactly as
 * sdbf::to_string writes each one (sdbf/sdbf_class.cc:294-334) and
 * sdbf_set::to_string concatenates them (sdbf/sdbf_set.cc:175-184); the base64 is
 * encoded on the device.  *len receives the length in bytes (no terminating
 * NUL).  out == NULL only sets *len; a cap below it is an error.
 */
int sdbfx_set_to_text(sdbfx_ctx *ctx, const sdbfx_set *set, char *out, uint64_t cap, uint64_t *len);

/*
 * Appends every digest of a text (lines in sdbf::to_string's form) to the
 * set: the reading loop of sdbf_set::sdbf_set(const char*) (sdbf/sdbf_set.cc:
 * 51-74, lines up to the first empty one) with each line read as
 * sdbf::sdbf(const std::string&) reads it (sdbf/sdbf_class.cc:144-215).  The
 * headers are read on the host; the base64 bodies and block-mode element
 * counts are decoded on the device straight into the set, whose Hamming
 * weights are computed there too.  Returns -1 for a line under 344 bytes and
 * -2 for a malformed one (the reference's throw -1 / throw -2), with nothing
 * added; on success *added (may be NULL) receives the number of digests.
 * Only 256-byte filters are accepted.
 */
int sdbfx_set_add_text(sdbfx_ctx *ctx, sdbfx_set *set, const char *text, uint64_t len, uint32_t *added);

/*
 * Fast mode: folds every digest of the set to 64-byte filters in place, as
 * sdbf::fast does for each item (sdbf/sdbf_class.cc:426-439, bloom_filter::
 * fold(2) at sdbf/bloom_filter.cc:263-289): each filter's four 64-byte
 * quarters are OR-ed into its first 64 bytes (the other 192 bytes stay) and
 * its Hamming weight becomes their popcount.  Folding again is a no-op;
 * digests added later stay plain until folded.  Replaces the fast() loops of
 * sdbf_set::compare_all (sdbf/sdbf_set.cc:243-248).
 *
 * A set may hold folded and plain digests.  As in sdbf_max_score
 * (sdbf/sdbf_core.cc:422-437), each oriented pair is scored by its
 * reference digest's mode: a folded reference digest on 64 bytes with
 * CUTOFFS64 and bf_bitcount_cut_64, against the target's filters and Hamming
 * weights as they are (folded or not); a plain reference digest on 256
 * bytes, also against a folded target, where the reference's formula uses
 * the target's folded Hamming weight (scores above 100 can result).  A pair
 * whose score is infinite there (the denominator max_est - cut_off is 0) makes
 * the comparison return -1, since the reference aborts on it
 * (boost::math::round raises on infinity).
 */
int sdbfx_set_fold(sdbfx_ctx *ctx, sdbfx_set *set);

/* Folds digest idx only (sdbf::fast on one item, sdbf/sdbf_class.cc:426-439). */
int sdbfx_set_fold_digest(sdbfx_ctx *ctx, sdbfx_set *set, uint32_t idx);

/* 1 when every digest of a non-empty set is folded, else 0. */
int sdbfx_set_folded(const sdbfx_set *set);

/* 1 when digest idx is folded (sdbf::fastmode), else 0. */
int sdbfx_set_digest_folded(const sdbfx_set *set, uint32_t idx);

/* ---- comparison ----------------------------------------------------- */

/*
 * All pairs i < j of `set`, in set order: scores[k] for the k-th pair in
 * row-major i<j order is items[i].compare(items[j], 0), an integer in
 * -1..100.  Replaces sdbf_set::compare_all / compare_all_quiet
 * (sdbf/sdbf_set.cc:240-308) -> sdbf::compare -> sdbf_score ->
 * sdbf_max_score (sdbf/sdbf_core.cc:331-455).  scores holds n*(n-1)/2 ints.
 */
int sdbfx_compare_all(sdbfx_ctx *ctx, const sdbfx_set *set, int32_t *scores);

/*
 * Query x target: scores[i*nt + j] = q[i].compare(t[j], 0).  Replaces
 * sdbf_set::compare_to / compare_to_quiet (sdbf/sdbf_set.cc:318-394).
 */
int sdbfx_compare_to(sdbfx_ctx *ctx, const sdbfx_set *q, const sdbfx_set *t, int32_t *scores);

/* As sdbfx_compare_to for the query digests [q_first, q_first + q_count)
 * only: scores[(i - q_first) * nt + j]. */
int sdbfx_compare_to_range(sdbfx_ctx *ctx, const sdbfx_set *q, uint32_t q_first, uint32_t q_count,
                           const sdbfx_set *t, int32_t *scores);

/*
 * As sdbfx_compare_to with sdbf_score's `sample` (sdhash -s N,
 * sdbf/sdbf_core.cc:342-374, via sdbf_set::compare_to(other, t, sample, f)):
 * a query digest with more than `sample` filters takes part with its first
 * `sample` filters, and the swap rule sees that count and the element count
 * of filter sample - 1.  The reference's random offsets (:370-372) are never
 * drawn, because after the swap rule the reference side never has more than
 * `sample` filters, so the result is deterministic.  sample 0 is
 * sdbfx_compare_to.
 */
int sdbfx_compare_to_sampled(sdbfx_ctx *ctx, const sdbfx_set *q, const sdbfx_set *t, uint32_t sample,
                             int32_t *scores);

/*
 * Selects the comparison kernel for later sdbfx_compare_* calls of the
 * context.  Both give identical scores:
 *   SDBFX_CMP_VALU  AND + popcount on the vector ALUs, one reference filter
 *                   per lane;
 *   SDBFX_CMP_MFMA  exact matrix-core products of the filter bits as FP4
 *                   (E2M1) elements 0 / 0.5 with block scale 2, f32
 *                   accumulation, 128 x 128 filter pairs per wave;
 *   SDBFX_CMP_MFMA_I8  the same tiling with int8 {0,1} byte planes and
 *                   int32 accumulation (half the FP4 form's rate);
 *   SDBFX_CMP_AUTO  (default) picks VALU or MFMA from the shape of the call.
 * No counterpart in the reference, whose bf_bitcount_cut_256
 * (sdbf/bf_utils.cc:115-206) is a single CPU loop.
 */
#define SDBFX_CMP_AUTO 0
#define SDBFX_CMP_VALU 1
#define SDBFX_CMP_MFMA 2
#define SDBFX_CMP_MFMA_I8 3
int sdbfx_set_compare_method(sdbfx_ctx *ctx, int method);

/*
 * Autotuning of the comparison launches (no counterpart in the reference,
 * whose comparison has a single CPU form).  With autotuning on and the
 * method AUTO, a call whose shape class (log2 of its reference-filter
 * instances, of its largest target digest's filter count and of its target
 * digest count) is new runs every candidate configuration on the call
 * itself, twice each -- the VALU kernel, and the MFMA kernel with 1024 to
 * 16384 target filters per work item -- checks that their scores agree, returns them, and
 * keeps the fastest for later calls of the class.  Without it (the default)
 * the engine's cost model and chunk rule decide.
 */
int sdbfx_set_autotune(sdbfx_ctx *ctx, int on);
typedef struct {
    uint64_t instances;        /* reference-filter instances of the tuning call */
    uint64_t max_target;       /* filters of its largest target digest */
    uint32_t target_digests;
    int32_t method;            /* SDBFX_CMP_VALU or SDBFX_CMP_MFMA */
    uint32_t chunk;            /* target filters per work item */
    float ms;                  /* the chosen configuration's time on that call */
} sdbfx_tune_entry;
/* The tuned classes so far: *n gets their number, entries[0 .. min(cap, n)) the entries. */
int sdbfx_autotune_table(const sdbfx_ctx *ctx, sdbfx_tune_entry *entries, uint32_t cap, uint32_t *n);

/* ---- serving -------------------------------------------------------- */

/* ---- indexes (sdhash --index / --index-search) ------------------------- */

/*
 * An index is bloom_filter(size, 5, 0, 0.01) (sdbf/bloom_filter.cc:52-71) in
 * device memory: `size` bytes (a power of two from 64 to 256 MiB; sdhash uses
 * 64 MiB, sdhash-src/sdhash_threads.cc:409,458), bit mask 8 * size - 1, the
 * five SHA-1 words of a feature as its positions, and an element count (the
 * inserts that set a clear bit, bloom_filter.cc:334-359).  Created empty.
 * Block-mode insertion also holds 32 bytes per index byte of scratch (2 GiB
 * for 64 MiB), allocated on its first use and freed with the index.
 */
int sdbfx_index_create(sdbfx_ctx *ctx, uint64_t size, sdbfx_index **out);
void sdbfx_index_destroy(sdbfx_index *ix);
uint64_t sdbfx_index_size(const sdbfx_index *ix);
/* Copies the index bits (size bytes; bits may be NULL) and the element count
 * to the host: the content of bloom_filter::write_out's .idx file
 * (bloom_filter.cc:208-228) before its LZ4 compression. */
int sdbfx_index_read(sdbfx_ctx *ctx, const sdbfx_index *ix, uint8_t *bits, uint64_t *elem_count);
/* Replaces the index content, e.g. with a .idx file's decompressed bits
 * (bloom_filter(string), bloom_filter.cc:77-114). */
int sdbfx_index_write(sdbfx_ctx *ctx, sdbfx_index *ix, const uint8_t *bits, uint64_t elem_count);

/*
 * sdbfx_set_digest with an index_info (sdbf/index_info.h), as the sdbf
 * constructors take it (sdbf_class.cc:71-136):
 *   - index (info->index, may be NULL): a stream-mode input skips every
 *     feature whose five index bits were all set, after setting its filter
 *     bits (gen_chunk_hash, sdbf_core.cc:110-115); a block-mode input inserts
 *     every counted feature (gen_block_hash, :163-166).  Either way the index
 *     takes the inserts.  Inputs are taken in the order given, as sdhash does
 *     with -p 1 (with more threads the reference's result depends on their
 *     timing), so the digests of stream inputs depend on everything inserted
 *     before them.
 *   - search[0 .. nsearch) (info->setlist's indexes, --index-search): in
 *     block mode, every fourth counted feature (k % 4 == 0 in counted order)
 *     is looked up in each (check_indexes, :167-177, 482-493) and nothing is
 *     inserted.  matches receives nsearch counts per dd-block of the call's
 *     block-mode inputs that are digested (inputs in order, blocks in order,
 *     block-major); matches_cap (in blocks) must hold them all.
 *     print_indexes (:458-472) reports a block when a count exceeds 4.
 * The digests are those of sdbfx_set_digest except for the stream gate.  A
 * call that fails may leave the inserts of its completed passes in the index.
 */
int sdbfx_set_digest_indexed(sdbfx_ctx *ctx, sdbfx_set *set, uint32_t n, const uint8_t *const *bufs,
                             const uint64_t *lens, const char *const *names, const uint32_t *dd_block,
                             sdbfx_index *index, const sdbfx_index *const *search, uint32_t nsearch,
                             uint32_t *matches, uint64_t matches_cap, int32_t *status);
/* As sdbfx_set_digest_indexed for inputs in device memory (as
 * sdbfx_set_digest_device). */
int sdbfx_set_digest_device_indexed(sdbfx_ctx *ctx, sdbfx_set *set, uint32_t n, const void *d_base,
                                    const uint64_t *offsets, const uint64_t *lens, const char *const *names,
                                    const uint32_t *dd_block, sdbfx_index *index, const sdbfx_index *const *search,
                                    uint32_t nsearch, uint32_t *matches, uint64_t matches_cap, int32_t *status);

/*
 * A query server over a resident reference set (continuous batching).
 * Queries submitted while earlier ones are on the GPU are queued and digested
 * together as the next batch (arrival order, at most batch_bytes of input and
 * batch_max queries per batch; a larger query goes alone); batch k+1 is
 * digested on one HIP stream while batch k is compared with the reference
 * set on another.  The server owns two contexts on `device`; `ref` must stay
 * alive and unchanged while the server exists.  Each result is
 * q.compare(ref[j]) for every digest j of `ref`, the same whether the query
 * was batched with others or alone.  Replaces the per-request hashString /
 * compareTwo path of sdhash-server (sdhash-srv-main.cc:224-402).
 */
typedef struct sdbfx_server sdbfx_server;
/* mem_budget (bytes, 0 = no limit) bounds the server's scratch device
 * memory, half for digesting and half for comparing: a batch that would
 * exceed it is split and its later queries deferred to the next batch; a
 * query that does not fit on its own is refused (status SDBFX_EBUDGET). */
int sdbfx_serve_create(int device, const sdbfx_set *ref, uint64_t batch_bytes, uint32_t batch_max,
                       uint64_t mem_budget, sdbfx_server **out);
/* Waits for the batches already past the queue; queued requests are dropped. */
void sdbfx_serve_destroy(sdbfx_server *s);
const char *sdbfx_serve_error(sdbfx_server *s);
uint32_t sdbfx_serve_ref_size(const sdbfx_server *s);
/* Copies the query; dd_block as in sdbfx_set_digest (0 = stream mode). */
int sdbfx_serve_submit(sdbfx_server *s, const uint8_t *buf, uint64_t len, const char *name, uint32_t dd_block,
                       uint64_t *ticket);
/* *done = 1 when the result is ready; then scores (sdbfx_serve_ref_size
 * ints), status (0, -3 for a query under 512 bytes, SDBFX_EBUDGET for one
 * refused by the memory budget, -1 on a device error) and the
 * submit-to-result latency are written and the ticket is released.  Any
 * output pointer may be NULL. */
int sdbfx_serve_poll(sdbfx_server *s, uint64_t ticket, int32_t *scores, int32_t *status, uint64_t *latency_ns,
                     int *done);
/* As poll, blocking until the result is ready. */
int sdbfx_serve_wait(sdbfx_server *s, uint64_t ticket, int32_t *scores, int32_t *status, uint64_t *latency_ns);
/* Batches and queries completed so far, and the batches compared with the
 * MFMA kernel.  The server picks the comparison kernel per batch size class
 * by timing both on the first batch of each class. */
int sdbfx_serve_stats(sdbfx_server *s, uint64_t *batches, uint64_t *queries, uint64_t *mfma_batches);
/* Queries deferred to a later batch and refused, by the memory budget. */
int sdbfx_serve_admission(sdbfx_server *s, uint64_t *deferred, uint64_t *refused);

/* ---- device memory for device-resident inputs ------------------------ */

/* Allocates `bytes` of device memory plus 256 readable padding bytes. */
int sdbfx_device_alloc(sdbfx_ctx *ctx, uint64_t bytes, void **dptr);
int sdbfx_device_free(sdbfx_ctx *ctx, void *dptr);
int sdbfx_upload(sdbfx_ctx *ctx, void *dst, const void *src, uint64_t bytes);

/* ---- no device allocation ------------------------------------------- */

typedef struct sdbfx_graph sdbfx_graph;

/*
 * Block-mode digest of one input already in device memory, into device
 * memory the caller provides, with no device allocation: no hipMalloc or
 * hipFree and no allocating library routine, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid it).  No counterpart in the reference.  The result equals
 * sdbfx_set_digest_device of that input with dd_block (the digest of
 * sdbf(name, buf, dd_block, len, NULL), sdbf/sdbf_class.cc:115-136):
 * d_filters receives *nfilters filters of 256 bytes, d_elem_counts and
 * d_hamming *nfilters values each (the caller sizes them for
 * len / dd_block + 1).  d_input must be 256-byte aligned with 256 readable
 * bytes past len; dd_block a multiple of 16 in [512, 16384]; inputs of more
 * than 131,072 dd-blocks are refused (-1), inputs under 512 bytes give -3.
 * The work list lives in a fixed device array of 3 MiB that every call on
 * the GPU shares: calls that overlap on one GPU race.  The kernels run on
 * the context's stream; the outputs are ready after sdbfx_sync.
 */
int sdbfx_digest_blocks_noalloc(sdbfx_ctx *ctx, const void *d_input, uint64_t len, uint32_t dd_block,
                                uint8_t *d_filters, uint16_t *d_elem_counts, uint16_t *d_hamming,
                                uint32_t *nfilters);

/*
 * The launches of sdbfx_digest_blocks_noalloc for one (d_input, len,
 * dd_block, outputs), captured once into a HIP graph on the context's stream
 * (no counterpart in the reference).  sdbfx_graph_launch replays it on the
 * context's stream: a loop that digests successive contents of one device
 * buffer (refilled by the caller in stream order between replays) then costs
 * one graph launch per call and no host work.  The graph holds no memset and
 * allocates nothing; it shares the no-allocation entry point's scratch array
 * (replays that overlap another no-allocation call on the GPU race with it)
 * and its limits.  *nfilters gets the filter count of every replay.
 */
int sdbfx_graph_blocks_create(sdbfx_ctx *ctx, const void *d_input, uint64_t len, uint32_t dd_block,
                              uint8_t *d_filters, uint16_t *d_elem_counts, uint16_t *d_hamming, uint32_t *nfilters,
                              sdbfx_graph **out);
int sdbfx_graph_launch(sdbfx_ctx *ctx, sdbfx_graph *g);
void sdbfx_graph_destroy(sdbfx_graph *g);

/* ---- kernel timing -------------------------------------------------- */

/* When enabled, every kernel launch of the context is bracketed by HIP
 * events on the context's stream.  sdbfx_kernel_time returns the summed
 * duration (ms) and launch count of kernel `name` since the last reset. */
int sdbfx_set_profiling(sdbfx_ctx *ctx, int on);
int sdbfx_kernel_time(sdbfx_ctx *ctx, const char *name, double *ms, uint64_t *launches);
int sdbfx_reset_kernel_times(sdbfx_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif

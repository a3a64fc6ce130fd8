/*
 * sdbf_oracle.c -- CPU restatement of the sdbf digest + compare hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (sdhash_amd/, include/)
 * links, imports or calls this file.  It is used by tests/ as the parity
 * checker, by __graft_entry__.smoke() as the checker of one small GPU call,
 * and by bench.py's cpu_baseline leg as the timed CPU port.
 *
 * Each function restates one reference function, literally and in the same
 * operation order, and cites the reference file:line it follows (paths are
 * relative to the sdhash source tree).  SHA-1 is taken from OpenSSL's
 * libcrypto, the library the reference calls (sdbf/sdbf_core.cc:105,160).
 *
 * Pinning status of each part (how it is checked independently of this file):
 *   entropy table        fingerprint of the reference build's table + formula
 *   ranks                closed form: direct 64-byte histogram per window
 *   scores state machine hand-traced known-answer vectors + closed forms on
 *                        monotone / constant / zero rank sequences
 *   bf_sha1_insert       set semantics (brute force) + FIPS-180 SHA-1 vectors
 *   stream hashing       hand-derived known answers over chosen scores and
 *                        SHA-1 words (orc_stream_hash_scored): zero-new-bit
 *                        skip, big-filter rejection keeping the filter bits,
 *                        rollover at 160, restart at 8738, the per-unit
 *                        restart count, the big filter kept across units,
 *                        the trim
 *   block hashing        hand-derived known answers (orc_block_digest_scored,
 *                        orc_block_hash_scored): allowance spent only on
 *                        counted features, the stop at 192, the threshold
 *                        walk with its k = 15 exit, the tail rule
 *   index gate / insert  hand-derived known answers (the *_indexed entries):
 *   / search             a feature already in the index is skipped by the
 *                        stream path after setting its filter bits; one with
 *                        no new filter bit never reaches the index; block mode
 *                        inserts every counted feature and queries every
 *                        fourth one; bloom_filter mask and insert/query by
 *                        brute force
 *   bitcount_cut         brute-force popcount
 *   fast mode            fold = OR of the four 64-byte quarters (identity of
 *                        the two halving folds); 64-byte counts by brute
 *                        force; self-compare = 100 and disjoint = 0 closed
 *                        forms; the CUTOFFS64 table fingerprint
 *   sdbf_score           closed forms (self-compare = 100, disjoint = 0,
 *                        all-sparse target = -1) + hand-computed pairs: the
 *                        swap on filter count, on last element count OR name,
 *                        the -1 sentinel's order dependence, the single-filter
 *                        denominator, rounding half up
 *   (tests/test_oracle_mutants.py plants one plausible mistake at a time in
 *   a copy of this file; the known-answer tests must fail on each)
 *   base64 / to_string   RFC 4648 vectors + format grammar
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdio.h>
#include <openssl/sha.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../sdhash_amd/csrc/sdbf_tables.h"

#define ORC_WIN 64             /* entr_win_size / pop_win_size (sdbf_conf.cc:22,25) */
#define ORC_SYNC 4096          /* block_size used for the entropy resync (sdbf_conf.cc:24) */
#define ORC_THRESHOLD 16       /* sdbf_conf.cc:29 */
#define ORC_MAX_ELEM 160       /* sdbf_defines.h:34 */
#define ORC_MAX_ELEM_DD 192    /* sdbf_defines.h:35 */
#define ORC_MIN_FILE 512       /* sdbf_defines.h:56 */
#define ORC_MIN_ELEM 16        /* sdbf_defines.h:58 */
#define ORC_BF 256             /* BF_SIZE, sdbf_defines.h:50 */
#define ORC_BIG 16384          /* BIGFILTER, sdbf_defines.h:64 */
#define ORC_BIG_ELEM 8738      /* BIGFILTER_ELEM, sdbf_defines.h:65 */
#define ORC_CHUNK (32ull << 20) /* stream chunk, sdbf_class.cc:123 */

static uint64_t ENTROPY[65];

/* sdbf/entr64.cc:14-21 (sdbf_conf::entr64_table_init_int) */
void orc_entropy_table(uint64_t *out)
{
    ENTROPY[0] = 0;
    for (uint32_t i = 1; i <= 64; i++) {
        double p = (double)i / 64;
        ENTROPY[i] = (uint64_t)((-p * (log(p) / log((double)2)) / 6) * 1024000);
    }
    if (out)
        memcpy(out, ENTROPY, sizeof(ENTROPY));
}

__attribute__((constructor)) static void orc_init(void) { orc_entropy_table(NULL); }

/* sdbf/entr64.cc:28-41 */
static uint64_t entr64_init(const uint8_t *buffer, uint8_t *ascii)
{
    memset(ascii, 0, 256);
    for (int i = 0; i < 64; i++)
        ascii[buffer[i]]++;
    uint64_t entr = 0;
    for (int i = 0; i < 256; i++)
        if (ascii[i])
            entr += ENTROPY[ascii[i]];
    return entr;
}

/* sdbf/entr64.cc:48-71 */
static uint64_t entr64_inc(uint64_t prev, const uint8_t *buffer, uint8_t *ascii)
{
    if (buffer[0] == buffer[64])
        return prev;
    uint32_t old_cnt = ascii[buffer[0]];
    uint32_t new_cnt = ascii[buffer[64]];
    ascii[buffer[0]]--;
    ascii[buffer[64]]++;
    if (old_cnt == new_cnt + 1)
        return prev;
    int64_t old_diff = (int64_t)ENTROPY[old_cnt] - (int64_t)ENTROPY[old_cnt - 1];
    int64_t new_diff = (int64_t)ENTROPY[new_cnt + 1] - (int64_t)ENTROPY[new_cnt];
    int64_t e = (int64_t)prev - old_diff + new_diff;
    if (e < 0)
        e = 0;
    else if (e > 1024000)
        e = 1024000;
    return (uint64_t)e;
}

/* sdbf/sdbf_core.cc:22-45 (gen_chunk_ranks, carryover 0) */
void orc_gen_ranks(const uint8_t *buf, uint64_t n, uint16_t *ranks)
{
    uint8_t ascii[256];
    uint64_t entropy = 0;
    memset(ranks, 0, n * sizeof(uint16_t));
    int64_t limit = (int64_t)n - ORC_WIN;
    for (int64_t off = 0; off < limit; off++) {
        if (off % ORC_SYNC == 0)
            entropy = entr64_init(buf + off, ascii);
        else
            entropy = entr64_inc(entropy, buf + off - 1, ascii);
        ranks[off] = SDBFX_ENTR64_RANKS[entropy >> 10];
    }
}

/* sdbf/sdbf_core.cc:51-89 (gen_chunk_scores) */
void orc_gen_scores(const uint16_t *r, uint64_t n, uint16_t *scores, int32_t *histo)
{
    uint64_t i, j;
    const uint64_t w = ORC_WIN;
    uint64_t min_pos = 0;
    uint16_t min_rank = r[min_pos];
    memset(scores, 0, n * sizeof(uint16_t));
    if (n <= w)
        return;
    for (i = 0; i < n - w; i++) {
        if (i > 0 && min_rank > 0) {
            while (r[i + w] >= min_rank && i < min_pos && i < n - w + 1) {
                if (r[i + w] == min_rank)
                    min_pos = i + w;
                scores[min_pos]++;
                i++;
            }
        }
        min_pos = i;
        min_rank = r[min_pos];
        for (j = i + 1; j < i + w; j++) {
            if (r[j] < min_rank && r[j]) {
                min_rank = r[j];
                min_pos = j;
            } else if (min_pos == j - 1 && r[j] == min_rank) {
                min_pos = j;
            }
        }
        if (r[min_pos] > 0)
            scores[min_pos]++;
    }
    if (histo)
        for (i = 0; i < n - w; i++)
            histo[scores[i]]++;
}

/* sdbf/bf_utils.cc:32-43 (bf_sha1_insert, bf_class 0 -> mask 0x7FF) */
uint32_t orc_bf_sha1_insert(uint8_t *bf, const uint32_t *h)
{
    uint32_t cnt = 0;
    for (int i = 0; i < 5; i++) {
        uint32_t ins = h[i] & 0x7FF;
        uint32_t k = ins >> 3;
        uint8_t bit = (uint8_t)(1u << (ins & 7));
        if (!(bf[k] & bit))
            cnt++;
        bf[k] |= bit;
    }
    return cnt;
}

/* sdbf/bloom_filter.cc:334-359 (query_and_set, mode_set, mask 0x1FFFF for the
 * 16 KiB filter built at sdbf_class.cc:379 -- bloom_filter.cc:59-66) */
static int big_insert(uint8_t *big, const uint32_t *h)
{
    uint32_t set = 0;
    for (int i = 0; i < 5; i++) {
        uint32_t pos = h[i] & 0x1FFFF;
        uint32_t k = pos >> 3;
        uint8_t bit = (uint8_t)(1u << (pos & 7));
        if (big[k] & bit)
            set++;
        else
            big[k] |= bit;
    }
    return set < 5;
}

/* An index: bloom_filter(size, 5, 0, 0.01) (sdbf/bloom_filter.cc:52-71) --
 * `size` bytes, a power of two >= 64, bit_mask = BIT_MASKS_32[log_size + 1]
 * with log_size the bit length of size (so bit_mask = 8 * size - 1), and
 * its element count. */
typedef struct {
    uint8_t *bf;
    uint64_t size;
    uint32_t mask;
    uint64_t elem_count;
} orc_index;

uint32_t orc_index_mask(uint64_t size)
{
    if (!size || (size & (size - 1)) || size < 64)
        return 0;
    uint32_t log_size = 0;
    for (uint64_t t = size; t; t >>= 1)
        log_size++;
    uint32_t bits = log_size + 2;                 /* BIT_MASKS_32[log_size + 1] has log_size + 2 ones */
    return bits >= 32 ? 0xFFFFFFFFu : ((1u << bits) - 1);
}

/* sdbf/bloom_filter.cc:334-359 (query_and_set): insert (mode_set) returns
 * whether some bit was clear, and counts the element; query returns whether
 * all five bits are set */
static int index_query_and_set(orc_index *ix, const uint32_t *h, int mode_set)
{
    uint32_t bit_cnt = 0;
    for (int i = 0; i < 5; i++) {
        uint32_t pos = h[i] & ix->mask;
        uint64_t k = pos >> 3;
        uint8_t bit = (uint8_t)(1u << (pos & 7));
        if (ix->bf[k] & bit) {
            bit_cnt++;
        } else {
            if (mode_set)
                ix->bf[k] |= bit;
            else
                return 0;
        }
    }
    if (mode_set) {
        if (bit_cnt < 5) {
            ix->elem_count++;
            return 1;
        }
        return 0;
    }
    return bit_cnt == 5;
}

/* What an sdbf's index_info supplies to the hashing (sdbf/index_info.h):
 * `index` (info->index: the stream-mode gate, sdbf_core.cc:110-115, and the
 * block-mode insert, :163-166) and the indexes of info->setlist
 * (`search`, queried by check_indexes, :167-177, 482-493); `matches`
 * receives per block the nsearch counts of check_indexes. */
typedef struct {
    orc_index *index;
    orc_index *const *search;
    uint32_t nsearch;
    uint32_t *matches;
} orc_info;

/* the 20 digest bytes reinterpreted as five host (little-endian) words,
 * as done by the (uint32_t*) casts at sdbf_core.cc:105-106 */
void orc_sha1_words(const uint8_t *p, uint32_t *h)
{
    uint8_t d[20];
    SHA1(p, 64, d);
    memcpy(h, d, 20);
}

/* ---------------------------------------------------------------- stream */

/* Where the SHA-1 words of the feature at a position come from: the input
 * bytes (the digest path), or a table of five words per position (the
 * known-answer tests, which choose each feature's bit positions by hand). */
typedef struct {
    const uint8_t *bytes;     /* input; the feature at i is bytes[i..i+63] */
    const uint32_t *words;    /* or: words[5*i .. 5*i+4] */
} feat_src;

static void feat_words(const feat_src *fs, uint64_t i, uint32_t *h)
{
    if (fs->words)
        memcpy(h, fs->words + 5 * i, 20);
    else
        orc_sha1_words(fs->bytes + i, h);
}

typedef struct {
    uint8_t *bf;
    uint32_t cap;        /* filters allocated */
    uint32_t bf_count;
    uint32_t last_count;
    uint8_t *big;        /* current big filter */
} stream_state;

/* sdbf/sdbf_core.cc:94-136 (gen_chunk_hash) over one unit: positions
 * [0, chunk_size - 64) of the unit that starts at `pos` of the feature
 * source; `index` (info->index, or NULL) gates the features that set a new
 * filter bit (:110-115) */
static void chunk_hash(stream_state *st, const feat_src *fs, uint64_t pos,
                       const uint16_t *scores, uint64_t chunk_size, orc_index *index)
{
    uint32_t h[5];
    uint32_t bf_count = st->bf_count, last_count = st->last_count;
    uint8_t *cur = st->bf + (uint64_t)(bf_count - 1) * ORC_BF;
    uint32_t bigfi_count = 0;
    if (chunk_size <= ORC_WIN)
        return;
    for (uint64_t i = 0; i < chunk_size - ORC_WIN; i++) {
        if (scores[i] > ORC_THRESHOLD) {
            feat_words(fs, pos + i, h);
            if (!orc_bf_sha1_insert(cur, h))
                continue;
            if (index && !index_query_and_set(index, h, 1))
                continue;
            if (!big_insert(st->big, h))
                continue;
            last_count++;
            bigfi_count++;
            if (last_count == ORC_MAX_ELEM) {
                cur += ORC_BF;
                bf_count++;
                last_count = 0;
            }
            if (bigfi_count == ORC_BIG_ELEM) {
                memset(st->big, 0, ORC_BIG);   /* new empty big filter pushed */
                bigfi_count = 0;
            }
        }
    }
    st->bf_count = bf_count;
    st->last_count = last_count;
}

/*
 * sdbf/sdbf_core.cc:197-246 (gen_chunk_sdbf): units of `chunk` bytes plus
 * the remainder, each scored (from the bytes' ranks, or taken from
 * `given` for the unit's positions) and hashed in order; then the final
 * trim.  The filter buffer holds ((len >> 11) + 1) filters, as the
 * reference's estimate (:202-204), plus one.
 */
static void stream_digest_core(const feat_src *fs, const uint16_t *given, uint64_t len, uint64_t chunk,
                               orc_index *index, uint8_t **filters_out, uint32_t *bf_count, uint32_t *last_count)
{
    uint64_t buff_size = ((len >> 11) + 1) << 8;
    if (buff_size < 256)
        buff_size = 256;
    stream_state st;
    st.cap = (uint32_t)(buff_size / ORC_BF);
    /* real scores credit at most one position per step, so features (score
     * > 16) are fewer than len / 17 and the estimate above holds; given
     * scores may make every position a feature */
    if (given)
        st.cap = (uint32_t)(len / ORC_MAX_ELEM + 2);
    st.bf = (uint8_t *)calloc(st.cap + 1, ORC_BF);
    st.bf_count = 1;
    st.last_count = 0;
    st.big = (uint8_t *)calloc(1, ORC_BIG);
    uint64_t nbuf = len < chunk ? len : chunk;
    uint16_t *ranks = given ? NULL : (uint16_t *)malloc(nbuf * sizeof(uint16_t));
    uint16_t *scores = given ? NULL : (uint16_t *)malloc(nbuf * sizeof(uint16_t));
    uint64_t qt = len / chunk, rem = len % chunk, pos = 0;
    int32_t histo[66];
    for (uint64_t i = 0; i < qt; i++, pos += chunk) {
        if (!given) {
            orc_gen_ranks(fs->bytes + chunk * i, chunk, ranks);
            memset(histo, 0, sizeof(histo));
            orc_gen_scores(ranks, chunk, scores, histo);
        }
        /* the threshold loop at sdbf_core.cc:220-224 has no effect in stream mode */
        chunk_hash(&st, fs, pos, given ? given + pos : scores, chunk, index);
    }
    if (rem > 0) {
        if (!given) {
            orc_gen_ranks(fs->bytes + qt * chunk, rem, ranks);
            orc_gen_scores(ranks, rem, scores, NULL);
        }
        chunk_hash(&st, fs, pos, given ? given + pos : scores, rem, index);
    }
    /* sdbf_core.cc:235-238 */
    if (st.bf_count > 1 && st.last_count < ORC_MAX_ELEM / 8) {
        st.bf_count--;
        st.last_count = ORC_MAX_ELEM;
    }
    free(ranks);
    free(scores);
    free(st.big);
    *filters_out = st.bf;
    *bf_count = st.bf_count;
    *last_count = st.last_count;
}

/*
 * Stream-mode digest of one buffer: sdbf_class.cc:115-136 (ctor from a
 * buffer, dd_block_size 0) -> sdbf_core.cc:197-246 (gen_chunk_sdbf).
 * Returns 0, or -3 when len < 512.  *filters_out is malloc'ed and holds
 * *bf_count filters (after the final trim).
 */
int orc_digest_stream(const uint8_t *buf, uint64_t len, uint8_t **filters_out,
                      uint32_t *bf_count, uint32_t *last_count)
{
    if (len < ORC_MIN_FILE)
        return -3;
    feat_src fs = {buf, NULL};
    stream_digest_core(&fs, NULL, len, ORC_CHUNK, NULL, filters_out, bf_count, last_count);
    return 0;
}

/*
 * Stream-mode digest with info->index set (the --index hashing of small
 * files, sdhash_threads.cc:376-431 -> sdbf_hash_files): index_bf holds the
 * index's index_size bytes and *index_elem its element count, both updated.
 * Inputs hashed one after another into the same index see each other's
 * features, so the caller runs them in order.
 */
int orc_digest_stream_indexed(const uint8_t *buf, uint64_t len, uint8_t *index_bf, uint64_t index_size,
                              uint64_t *index_elem, uint8_t **filters_out, uint32_t *bf_count, uint32_t *last_count)
{
    if (len < ORC_MIN_FILE)
        return -3;
    orc_index ix = {index_bf, index_size, orc_index_mask(index_size), *index_elem};
    feat_src fs = {buf, NULL};
    stream_digest_core(&fs, NULL, len, ORC_CHUNK, &ix, filters_out, bf_count, last_count);
    *index_elem = ix.elem_count;
    return 0;
}

/*
 * Known-answer entry of the stream path: the same unit loop, hashing and
 * trim as orc_digest_stream, with the per-position scores and SHA-1 words
 * given (scores[i] for every input position i, words[5*i..5*i+4]) and the
 * unit size a parameter (the digest path uses 32 MiB).
 */
void orc_stream_hash_scored(const uint16_t *scores, const uint32_t *words, uint64_t len, uint64_t unit,
                            uint8_t **filters_out, uint32_t *bf_count, uint32_t *last_count)
{
    feat_src fs = {NULL, words};
    stream_digest_core(&fs, scores, len, unit, NULL, filters_out, bf_count, last_count);
}

/* orc_stream_hash_scored with an index gate (index_bf / index_size /
 * *index_elem as orc_digest_stream_indexed) */
void orc_stream_hash_scored_indexed(const uint16_t *scores, const uint32_t *words, uint64_t len, uint64_t unit,
                                    uint8_t *index_bf, uint64_t index_size, uint64_t *index_elem,
                                    uint8_t **filters_out, uint32_t *bf_count, uint32_t *last_count)
{
    orc_index ix = {index_bf, index_size, orc_index_mask(index_size), *index_elem};
    feat_src fs = {NULL, words};
    stream_digest_core(&fs, scores, len, unit, &ix, filters_out, bf_count, last_count);
    *index_elem = ix.elem_count;
}

/* ---------------------------------------------------------------- block */

/* sdbf/sdbf_core.cc:142-191 (gen_block_hash): the positions of one block
 * start at `pos` of the feature source.  With info: without searched
 * indexes every counted feature goes into info->index (:163-166); with
 * them, every fourth counted feature (hash_cnt % 4 == 0 before the count)
 * is looked up in each (check_indexes, :167-177) and the block's hits per
 * index go to matches[0 .. nsearch).  (The hashes[] copies and their cap at
 * :169-175 only stage the lookups.) */
static uint32_t block_hash(const feat_src *fs, uint64_t pos, const uint16_t *scores, uint32_t max_offset,
                           uint32_t threshold, int32_t allowed, uint8_t *bf, const orc_info *info,
                           uint32_t *matches)
{
    uint32_t h[5], hash_cnt = 0;
    const uint32_t nsearch = info ? info->nsearch : 0;
    for (uint32_t m = 0; m < nsearch; m++)
        matches[m] = 0;
    for (uint32_t i = 0; i < max_offset - ORC_WIN && hash_cnt < ORC_MAX_ELEM_DD; i++) {
        if (scores[i] > threshold || (scores[i] == threshold && allowed > 0)) {
            feat_words(fs, pos + i, h);
            if (!orc_bf_sha1_insert(bf, h))
                continue;
            if (nsearch == 0) {
                if (info && info->index)
                    (void)index_query_and_set(info->index, h, 1);
            } else if (hash_cnt % 4 == 0) {
                for (uint32_t m = 0; m < nsearch; m++)
                    if (index_query_and_set(info->search[m], h, 0))
                        matches[m]++;
            }
            hash_cnt++;
            if (scores[i] == threshold)
                allowed--;
        }
    }
    return hash_cnt;
}

/* sdbf/sdbf_core.cc:271-276 -- threshold k and allowance from the histogram */
void orc_block_threshold(const int32_t *histo, uint32_t *k_out, int32_t *allowed_out)
{
    uint32_t k, sum;
    for (k = 65, sum = 0; k >= ORC_THRESHOLD; k--) {
        if ((sum <= ORC_MAX_ELEM_DD) && (sum + (uint32_t)histo[k] > ORC_MAX_ELEM_DD))
            break;
        sum += histo[k];
    }
    *k_out = k;
    *allowed_out = (int32_t)(ORC_MAX_ELEM_DD - sum);
}

uint64_t orc_block_count(uint64_t len, uint32_t dd)
{
    return len / dd + ((len % dd) >= ORC_MIN_FILE ? 1 : 0);
}

/*
 * Block-mode digest: sdbf_class.cc:124-133 -> sdbf_core.cc:288-324
 * (gen_block_sdbf_mt, with thread_gen_block_sdbf at :252-282 per full
 * block).  Each full block: ranks, scores with their histogram (or the
 * given scores and their histogram over [0, dd - 64), as
 * gen_chunk_scores counts it at :83-86), threshold and allowance, then
 * gen_block_hash.  The tail (remainder >= 512 bytes, :309-319): threshold
 * 16 and allowance 192.  filters holds orc_block_count() * 256 zeroed bytes,
 * elem_counts orc_block_count() entries; blocks are independent, so
 * `threads` > 1 spreads them over OpenMP threads.
 */
static void block_digest_core(const feat_src *fs, const uint16_t *given, uint64_t len, uint32_t dd,
                              uint8_t *filters, uint16_t *elem_counts, int threads, const orc_info *info)
{
    uint64_t qt = len / dd, rem = len % dd;
    /* with an index the blocks run in order (an index insert is not
     * thread-safe, and its element count depends on the order) */
    if (info)
        threads = 1;
#ifdef _OPENMP
    if (threads < 1)
        threads = 1;
#pragma omp parallel num_threads(threads)
#endif
    {
        uint16_t *ranks = (uint16_t *)malloc(dd * sizeof(uint16_t));
        uint16_t *scores = (uint16_t *)malloc(dd * sizeof(uint16_t));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t i = 0; i < (int64_t)qt; i++) {
            const uint64_t pos = (uint64_t)i * dd;
            int32_t histo[66];
            uint32_t k;
            int32_t allowed;
            const uint16_t *sc = scores;
            memset(histo, 0, sizeof(histo));
            if (given) {
                sc = given + pos;
                for (uint32_t j = 0; j < dd - ORC_WIN; j++)
                    histo[sc[j]]++;
            } else {
                orc_gen_ranks(fs->bytes + pos, dd, ranks);
                orc_gen_scores(ranks, dd, scores, histo);
            }
            orc_block_threshold(histo, &k, &allowed);
            elem_counts[i] = (uint16_t)block_hash(fs, pos, sc, dd, k, allowed, filters + (uint64_t)i * ORC_BF, info,
                                                  info ? info->matches + (uint64_t)i * info->nsearch : NULL);
        }
        free(ranks);
        free(scores);
    }
    (void)threads;
    if (rem >= ORC_MIN_FILE) {
        uint16_t *ranks = (uint16_t *)malloc(dd * sizeof(uint16_t));
        uint16_t *scores = (uint16_t *)malloc(dd * sizeof(uint16_t));
        const uint16_t *sc = scores;
        if (given) {
            sc = given + qt * dd;
        } else {
            orc_gen_ranks(fs->bytes + qt * dd, rem, ranks);
            orc_gen_scores(ranks, rem, scores, NULL);
        }
        elem_counts[qt] = (uint16_t)block_hash(fs, qt * dd, sc, (uint32_t)rem, ORC_THRESHOLD, ORC_MAX_ELEM_DD,
                                               filters + qt * ORC_BF, info,
                                               info ? info->matches + qt * info->nsearch : NULL);
        free(ranks);
        free(scores);
    }
}

int orc_digest_block(const uint8_t *buf, uint64_t len, uint32_t dd, uint8_t *filters,
                     uint16_t *elem_counts, int threads)
{
    if (len < ORC_MIN_FILE)
        return -3;
    feat_src fs = {buf, NULL};
    block_digest_core(&fs, NULL, len, dd, filters, elem_counts, threads, NULL);
    return 0;
}

/*
 * Block-mode digest with an index_info: index_bf (index_size bytes,
 * *index_elem elements; or NULL) is info->index, filled by every counted
 * feature (the --index hashing of large files, sdhash_threads.cc:432-489);
 * search_bfs / search_sizes (nsearch of them) are the indexes of
 * info->setlist (--index-search, sdhash.cc:241-264), and matches receives
 * orc_block_count() * nsearch counts, block-major.
 */
int orc_digest_block_indexed(const uint8_t *buf, uint64_t len, uint32_t dd, uint8_t *index_bf, uint64_t index_size,
                             uint64_t *index_elem, uint8_t *const *search_bfs, const uint64_t *search_sizes,
                             uint32_t nsearch, uint8_t *filters, uint16_t *elem_counts, uint32_t *matches)
{
    if (len < ORC_MIN_FILE)
        return -3;
    orc_index ix = {index_bf, index_size, index_bf ? orc_index_mask(index_size) : 0, index_elem ? *index_elem : 0};
    orc_index *sx = nsearch ? (orc_index *)calloc(nsearch, sizeof(orc_index)) : NULL;
    orc_index **sp = nsearch ? (orc_index **)calloc(nsearch, sizeof(orc_index *)) : NULL;
    for (uint32_t m = 0; m < nsearch; m++) {
        sx[m].bf = search_bfs[m];
        sx[m].size = search_sizes[m];
        sx[m].mask = orc_index_mask(search_sizes[m]);
        sp[m] = &sx[m];
    }
    orc_info info = {index_bf ? &ix : NULL, sp, nsearch, matches};
    feat_src fs = {buf, NULL};
    block_digest_core(&fs, NULL, len, dd, filters, elem_counts, 1, &info);
    if (index_elem)
        *index_elem = ix.elem_count;
    free(sx);
    free(sp);
    return 0;
}

/*
 * Known-answer entries of the block path, with per-position scores (each
 * <= 65) and SHA-1 words given: the whole digest of `len` positions
 * (block_digest_core as above), and one gen_block_hash call with its
 * threshold and allowance chosen by the caller (returns hash_cnt).
 */
void orc_block_digest_scored(const uint16_t *scores, const uint32_t *words, uint64_t len, uint32_t dd,
                             uint8_t *filters, uint16_t *elem_counts)
{
    feat_src fs = {NULL, words};
    block_digest_core(&fs, scores, len, dd, filters, elem_counts, 1, NULL);
}

uint32_t orc_block_hash_scored(const uint16_t *scores, const uint32_t *words, uint32_t max_offset,
                               uint32_t threshold, int32_t allowed, uint8_t *bf)
{
    feat_src fs = {NULL, words};
    return block_hash(&fs, 0, scores, max_offset, threshold, allowed, bf, NULL, NULL);
}

/* orc_block_hash_scored with an index_info: index_bf (or NULL) filled, or
 * the nsearch indexes queried with the counts in matches[0 .. nsearch) */
uint32_t orc_block_hash_scored_indexed(const uint16_t *scores, const uint32_t *words, uint32_t max_offset,
                                       uint32_t threshold, int32_t allowed, uint8_t *bf, uint8_t *index_bf,
                                       uint64_t index_size, uint64_t *index_elem, uint8_t *const *search_bfs,
                                       const uint64_t *search_sizes, uint32_t nsearch, uint32_t *matches)
{
    orc_index ix = {index_bf, index_size, index_bf ? orc_index_mask(index_size) : 0, index_elem ? *index_elem : 0};
    orc_index sx[64], *sp[64];
    if (nsearch > 64)
        return 0xFFFFFFFFu;
    for (uint32_t m = 0; m < nsearch; m++) {
        sx[m].bf = search_bfs[m];
        sx[m].size = search_sizes[m];
        sx[m].mask = orc_index_mask(search_sizes[m]);
        sx[m].elem_count = 0;
        sp[m] = &sx[m];
    }
    orc_info info = {index_bf ? &ix : NULL, sp, nsearch, matches};
    feat_src fs = {NULL, words};
    uint32_t r = block_hash(&fs, 0, scores, max_offset, threshold, allowed, bf, &info, matches);
    if (index_elem)
        *index_elem = ix.elem_count;
    return r;
}

/* bloom_filter::insert_sha1 / query_sha1 of one hash (five words) */
int orc_index_insert(uint8_t *index_bf, uint64_t index_size, uint64_t *index_elem, const uint32_t *h)
{
    orc_index ix = {index_bf, index_size, orc_index_mask(index_size), *index_elem};
    int r = index_query_and_set(&ix, h, 1);
    *index_elem = ix.elem_count;
    return r;
}

int orc_index_query(uint8_t *index_bf, uint64_t index_size, const uint32_t *h)
{
    orc_index ix = {index_bf, index_size, orc_index_mask(index_size), 0};
    return index_query_and_set(&ix, h, 0);
}

/* ------------------------------------------------------------- compare */

/* A digest as seen by the comparison (the fields sdbf_score reads). */
typedef struct {
    const uint8_t *bf;          /* bf_count * 256 bytes */
    const uint16_t *elem_counts;/* dd mode: per filter; NULL for stream mode */
    uint16_t *hamming;          /* filled by orc_hamming */
    const char *name;
    uint32_t bf_count;
    uint32_t last_count;        /* stream mode */
    uint32_t max_elem;          /* 160 stream, 192 dd */
    uint32_t fast;              /* filters folded to 64 bytes (sdbf::fast) */
} orc_digest;

/* sdbf/sdbf_class.cc:389-401 */
void orc_hamming(const uint8_t *bf, uint32_t bf_count, uint16_t *ham)
{
    for (uint32_t i = 0; i < bf_count; i++) {
        uint32_t c = 0;
        for (int b = 0; b < ORC_BF; b++)
            c += (uint32_t)__builtin_popcount(bf[(uint64_t)i * ORC_BF + b]);
        ham[i] = (uint16_t)c;
    }
}

/* sdbf/sdbf_class.cc:407-415 */
static int32_t elem_count(const orc_digest *d, uint64_t idx)
{
    if (!d->elem_counts)
        return (idx < (uint64_t)d->bf_count - 1) ? (int32_t)d->max_elem : (int32_t)d->last_count;
    return d->elem_counts[idx];
}

/* sdbf/bf_utils.cc:115-206 (bf_bitcount_cut_256_asm): AND-popcount with
 * early exits after 32, 64 and 128 bytes; cut_off 0 counts all 256 bytes */
uint32_t orc_bitcount_cut_256(const uint8_t *a, const uint8_t *b, uint32_t cut_off, int32_t slack)
{
    uint64_t x[32], y[32];
    memcpy(x, a, 256);
    memcpy(y, b, 256);
    uint32_t result = 0;
    for (int w = 0; w < 4; w++)
        result += (uint32_t)__builtin_popcountll(x[w] & y[w]);
    if (cut_off > 0 && (8 * result + slack) < cut_off)
        return 0;
    for (int w = 4; w < 8; w++)
        result += (uint32_t)__builtin_popcountll(x[w] & y[w]);
    if (cut_off > 0 && (4 * result + slack) < cut_off)
        return 0;
    for (int w = 8; w < 16; w++)
        result += (uint32_t)__builtin_popcountll(x[w] & y[w]);
    if (cut_off > 0 && (2 * result + slack) < cut_off)
        return 0;
    for (int w = 16; w < 32; w++)
        result += (uint32_t)__builtin_popcountll(x[w] & y[w]);
    return result;
}

/* sdbf/bf_utils.cc:79-110 (bf_bitcount_cut_64_asm): AND-popcount of 64-byte
 * (folded) filters with one early exit after 32 bytes */
uint32_t orc_bitcount_cut_64(const uint8_t *a, const uint8_t *b, uint32_t cut_off, int32_t slack)
{
    uint64_t x[8], y[8];
    memcpy(x, a, 64);
    memcpy(y, b, 64);
    uint32_t result = 0;
    for (int w = 0; w < 4; w++)
        result += (uint32_t)__builtin_popcountll(x[w] & y[w]);
    if (cut_off > 0 && (2 * result + slack) < cut_off)
        return 0;
    for (int w = 4; w < 8; w++)
        result += (uint32_t)__builtin_popcountll(x[w] & y[w]);
    return result;
}

/* sdbf/sdbf_class.cc:426-439 (sdbf::fast) with bloom_filter::fold(2)
 * (sdbf/bloom_filter.cc:263-289) and compute_hamming (:453-461): each
 * filter's 32 64-bit words are folded twice (w[j] |= w[j + 16] for j < 16,
 * then w[j] |= w[j + 8] for j < 8), the 64 folded bytes replace the filter's
 * first 64 bytes (the other 192 stay), and ham[i] becomes their popcount. */
void orc_fold_fast(uint8_t *bf, uint32_t bf_count, uint16_t *ham)
{
    for (uint32_t i = 0; i < bf_count; i++) {
        uint64_t w[32];
        memcpy(w, bf + (uint64_t)i * ORC_BF, ORC_BF);
        uint64_t rsize = 32;
        for (int t = 0; t < 2; t++) {
            for (uint64_t j = 0; j < rsize / 2; j++)
                w[j] |= w[j + rsize / 2];
            rsize /= 2;
        }
        memcpy(bf + (uint64_t)i * ORC_BF, w, 64);
        uint32_t c = 0;
        for (int j = 0; j < 8; j++)
            c += (uint32_t)__builtin_popcountll(w[j]);
        ham[i] = (uint16_t)c;
    }
}

/* sdbf/sdbf_core.cc:396-455 (sdbf_max_score): 256-byte filters, or the folded
 * 64-byte ones when the reference digest is in fast mode (:422-437) */
double orc_max_score(const orc_digest *ref, uint32_t ref_index, const orc_digest *tgt)
{
    double score, max_score = -1;
    uint32_t s1, s2, max_est, match, cut_off, slack = 48;
    s1 = (uint32_t)elem_count(ref, ref_index);
    if (s1 < ORC_MIN_ELEM)
        return 0;
    const uint8_t *bf_1 = ref->bf + (uint64_t)ref_index * ORC_BF;
    uint32_t e1 = ref->hamming[ref_index];
    for (uint32_t i = 0; i < tgt->bf_count; i++) {
        const uint8_t *bf_2 = tgt->bf + (uint64_t)i * ORC_BF;
        s2 = (uint32_t)elem_count(tgt, i);
        if (ref->bf_count >= 1 && s2 < ORC_MIN_ELEM)
            continue;
        uint32_t e2 = tgt->hamming[i];
        max_est = (e1 < e2) ? e1 : e2;
        if (!ref->fast) {
            uint32_t mn = 4096 / (s1 + s2);
            cut_off = SDBFX_CUTOFFS256[mn];
            match = orc_bitcount_cut_256(bf_1, bf_2, cut_off, slack);
            if (match > 0)
                match = orc_bitcount_cut_256(bf_1, bf_2, 0, 0);
        } else {
            uint32_t mn = 1024 / (s1 + s2);
            cut_off = SDBFX_CUTOFFS64[mn];
            match = orc_bitcount_cut_64(bf_1, bf_2, cut_off, slack);
            if (match > 0)
                match = orc_bitcount_cut_64(bf_1, bf_2, 0, 0);
        }
        score = (match <= cut_off) ? 0 : (double)(match - cut_off) / (max_est - cut_off);
        max_score = (score > max_score) ? score : max_score;
    }
    return max_score;
}

/* sdbf/sdbf_core.cc:331-389 (sdbf_score) with the boost 1.49 round of
 * external/boost/math/special_functions/round.hpp:20-26.  With sample > 0 a
 * calling digest of more than `sample` filters counts as `sample` filters in
 * the swap rule (:342-359) and is scored on its first `sample` filters: the
 * loop's random offset (:370-372) needs bf_count_1 > sample, which the code
 * before it never leaves true, so ref_index is always i. */
int orc_score_sampled(const orc_digest *a, const orc_digest *b, uint32_t sample)
{
    const orc_digest *s1 = a, *s2 = b;
    double score_sum = -1;
    uint32_t n1;
    if (sample > 0 && s1->bf_count > sample)
        n1 = sample;
    else
        n1 = s1->bf_count;
    if ((n1 > s2->bf_count) ||
        (n1 == s2->bf_count &&
         ((elem_count(s1, n1 - 1) > elem_count(s2, s2->bf_count - 1)) ||
          strcmp(s1->name, s2->name) > 0))) {
        const orc_digest *t = s1;
        s1 = s2;
        s2 = t;
        n1 = s1->bf_count;
    }
    uint32_t rand_offset = 1;
    int sparse = 0;
    for (uint32_t i = 0; i < n1; i++) {
        if (sample > 0 && n1 > sample)
            abort();   /* unreachable: see above */
        double m = orc_max_score(s1, i * rand_offset, s2);
        score_sum = (score_sum < 0) ? m : score_sum + m;
        if (elem_count(s1, i) < ORC_MIN_ELEM)
            sparse++;
    }
    uint64_t denom = n1;
    if (n1 > 1)
        denom -= sparse;
    if (denom == 0)
        score_sum = -1;
    if (score_sum < 0)
        return -1;
    double v = 100.0 * score_sum / (double)denom;
    /* an infinite fraction (a plain reference filter against a folded
     * target whose Hamming weight equals the cutoff: max_est - cut_off is 0)
     * makes boost::math::round raise in the reference; reported as INT_MIN */
    if (isinf(v))
        return (int)0x80000000u;
    return (int)floor(v + 0.5);
}

int orc_score(const orc_digest *a, const orc_digest *b) { return orc_score_sampled(a, b, 0); }

/*
 * All pairs i<j of a set, in set order (sdbf_set.cc:240-266): out[k] holds
 * items[i].compare(items[j]) for the k-th pair in row-major i<j order.
 */
void orc_compare_all(const orc_digest *set, uint32_t n, int32_t *out, int threads)
{
    (void)threads;
#ifdef _OPENMP
    if (threads < 1)
        threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
    for (int64_t i = 0; i < (int64_t)n; i++) {
        uint64_t base = (uint64_t)i * n - (uint64_t)i * (i + 1) / 2;
        for (uint32_t j = (uint32_t)i + 1; j < n; j++)
            out[base + (j - i - 1)] = orc_score(&set[i], &set[j]);
    }
}

/* query x target (sdbf_set.cc:318-348): out[i*nt + j] = q[i].compare(t[j], sample) */
void orc_compare_to_sampled(const orc_digest *q, uint32_t nq, const orc_digest *t, uint32_t nt,
                            uint32_t sample, int32_t *out, int threads)
{
    (void)threads;
#ifdef _OPENMP
    if (threads < 1)
        threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
    for (int64_t i = 0; i < (int64_t)nq; i++)
        for (uint32_t j = 0; j < nt; j++)
            out[(uint64_t)i * nt + j] = orc_score_sampled(&q[i], &t[j], sample);
}

void orc_compare_to(const orc_digest *q, uint32_t nq, const orc_digest *t, uint32_t nt,
                    int32_t *out, int threads)
{
    (void)threads;
#ifdef _OPENMP
    if (threads < 1)
        threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
    for (int64_t i = 0; i < (int64_t)nq; i++)
        for (uint32_t j = 0; j < nt; j++)
            out[(uint64_t)i * nt + j] = orc_score(&q[i], &t[j]);
}

/* Stream digests of many buffers, one per OpenMP task (the CPU baseline's
 * file-parallel hashing, sdhash_threads.cc:153-186 / 241-297). */
void orc_digest_stream_many(uint32_t n, const uint8_t *const *bufs, const uint64_t *lens,
                            uint8_t **filters, uint32_t *bf_counts, uint32_t *last_counts,
                            int32_t *status, int threads)
{
    (void)threads;
#ifdef _OPENMP
    if (threads < 1)
        threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
    for (int64_t i = 0; i < (int64_t)n; i++)
        status[i] = orc_digest_stream(bufs[i], lens[i], &filters[i], &bf_counts[i], &last_counts[i]);
}

void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------- base64 */

static const char B64[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

/* standard alphabet with '=' padding, as modp_b64_encode (base64/modp_b64.cc:79-118) */
uint64_t orc_b64encode(const uint8_t *in, uint64_t len, char *out)
{
    uint64_t i, o = 0;
    for (i = 0; i + 2 < len; i += 3) {
        uint32_t v = ((uint32_t)in[i] << 16) | ((uint32_t)in[i + 1] << 8) | in[i + 2];
        out[o++] = B64[(v >> 18) & 63];
        out[o++] = B64[(v >> 12) & 63];
        out[o++] = B64[(v >> 6) & 63];
        out[o++] = B64[v & 63];
    }
    if (len - i == 1) {
        uint32_t v = (uint32_t)in[i] << 16;
        out[o++] = B64[(v >> 18) & 63];
        out[o++] = B64[(v >> 12) & 63];
        out[o++] = '=';
        out[o++] = '=';
    } else if (len - i == 2) {
        uint32_t v = ((uint32_t)in[i] << 16) | ((uint32_t)in[i + 1] << 8);
        out[o++] = B64[(v >> 18) & 63];
        out[o++] = B64[(v >> 12) & 63];
        out[o++] = B64[(v >> 6) & 63];
        out[o++] = '=';
    }
    out[o] = 0;
    return o;
}

/*
 * sdbf::to_string (sdbf_class.cc:294-334).  Returns the number of bytes
 * written (no terminator counted) or the size needed when out is NULL.
 */
uint64_t orc_to_string(const orc_digest *d, uint64_t orig_size, uint32_t dd_block_size, char *out)
{
    uint64_t need = 256 + strlen(d->name) + (uint64_t)d->bf_count * 350 + 8;
    if (!out)
        return need;
    char *p = out;
    if (!d->elem_counts) {
        p += sprintf(p, "sdbf:03:%d:%s:%llu:sha1:256:5:7ff:%u:%u:%u:", (int)strlen(d->name), d->name,
                     (unsigned long long)orig_size, d->max_elem, d->bf_count, d->last_count);
        /* groups of 6 filters are 1536 bytes (a multiple of 3), so the
         * concatenated base64 equals one encoding of the whole buffer */
        p += orc_b64encode(d->bf, (uint64_t)d->bf_count * ORC_BF, p);
    } else {
        p += sprintf(p, "sdbf-dd:03:%d:%s:%llu:sha1:256:5:7ff:%u:%u:%u", (int)strlen(d->name), d->name,
                     (unsigned long long)orig_size, d->max_elem, d->bf_count, dd_block_size);
        for (uint32_t i = 0; i < d->bf_count; i++) {
            p += sprintf(p, ":%02x:", d->elem_counts[i]);
            p += orc_b64encode(d->bf + (uint64_t)i * ORC_BF, ORC_BF, p);
        }
    }
    *p++ = '\n';
    *p = 0;
    return (uint64_t)(p - out);
}

/* the per-filter element count of a digest (for tests) */
int32_t orc_elem_count(const orc_digest *d, uint64_t idx) { return elem_count(d, idx); }

int orc_threads_available(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

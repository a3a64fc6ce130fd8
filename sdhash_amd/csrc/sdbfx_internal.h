// Structures shared by the host scheduler (sdbfx_host.cpp) and the gfx950
// kernels (sdbfx_kernels.hip).  Plain data only.
#ifndef SDBFX_INTERNAL_H
#define SDBFX_INTERNAL_H
#include <stdint.h>

#define SDBFX_WIN 64             // entropy window == popularity window
#define SDBFX_BF 256             // filter bytes
#define SDBFX_BFW 64             // filter 32-bit words
#define SDBFX_MAX_ELEM 160       // stream-mode elements per filter
#define SDBFX_MAX_ELEM_DD 192    // block-mode elements per filter
#define SDBFX_MIN_ELEM 16        // filters with fewer elements are sparse
#define SDBFX_THRESHOLD 16       // popularity threshold
#define SDBFX_MIN_FILE 512       // smallest digestible input
#define SDBFX_BIG_BYTES 16384    // per-input de-duplication filter
#define SDBFX_BIG_ELEM 8738      // elements before the de-duplication filter restarts
#define SDBFX_CHUNK (32ull << 20) // stream-mode restart unit

// digest workgroups
#define SDBFX_NT 256             // threads per digest workgroup
#define SDBFX_TILE 16384         // owned score positions per stream tile
// 1: stream insertion replays every batch one candidate at a time (a test
// setting: the replay is otherwise taken only when a batch decides a
// rejection internally)
#ifndef SDBFX_INSERT_REPLAY
#define SDBFX_INSERT_REPLAY 0
#endif
#ifndef SDBFX_WARM
#define SDBFX_WARM 64            // speculative warm-up steps
#endif
#define SDBFX_CAND_CAP 1024      // candidate slots per stream tile (>= (TILE+64)/17)
#define SDBFX_PAD 256            // alignment / padding of inputs in device memory

// compare
#define SDBFX_CMP_CHUNK 1024     // target filters per compare work item (VALU kernel; the MFMA
                                 // kernel takes 1024-16384 per call)
#define SDBFX_CMP_ROWS 64        // reference-filter instances per work item, VALU compare
#define SDBFX_MROWS 128          // reference-filter instances per work item, MFMA compare
#define SDBFX_SCORE_NONFINITE ((int32_t)0x80000000)   // pair score whose sum is infinite

// One dd-block of block mode (also the tail block of an input).
struct BlockDesc {
    uint64_t off;       // byte offset of the block in the data buffer
    uint32_t n;         // block length (dd size, or the tail remainder >= 512)
    uint32_t tail;      // 1: tail block (threshold 16, allowance 192)
    uint64_t out;       // filter index in the output filter array
};

// One stream-mode tile: owned score positions [t0, t0+T) of a unit
// (a 32 MiB chunk of an input, or the input's remainder).
struct TileDesc {
    uint64_t unit_off;  // byte offset of the unit in the data buffer
    uint32_t n;         // unit length
    uint32_t t0;        // first owned position (unit-relative)
    uint32_t T;         // owned positions
    uint32_t first;     // 1: first tile of its unit
};

// A stream-mode input for the ordered insertion pass.
struct StreamFile {
    uint32_t tile_begin, tile_end;   // its tiles, in order
    uint64_t filt_base;              // first staging filter slot
    uint32_t filt_cap;               // staging filter slots
    uint32_t pad;
};

// Per-tile state written by the tile kernel.
struct TileState {
    uint32_t count;      // candidates (score > 16) in the tile
    uint32_t handoff;    // canonical state before the next tile's first iteration
    uint32_t certified;  // 1 when the tile's incoming state was proven exact
    uint32_t pad;
};

// A stream unit for the fix-up pass (tiles [tile_begin, tile_end)).
struct UnitDesc {
    uint32_t tile_begin, tile_end;
};

// Compare: all work items of one target digest.
struct CmpGroup {
    uint64_t tgt_first;    // first filter of the target digest
    uint32_t tgt_count;    // its filters
    uint32_t nch;          // target chunks
    uint64_t inst_begin;   // first reference-filter instance of the group
    uint32_t inst_count;   // instances
    uint32_t nblk;         // blocks of 64 instances
    uint32_t pair_begin;   // pairs of the group (sorted by instance offset)
    uint32_t pair_end;
    uint64_t wi_begin;     // first work item of the group
    uint64_t partial_begin;// first partial-maximum slot of the group
    uint32_t tgt_set;      // 0: target filters in set A, 1: in set B
    uint32_t chunk;        // target filters per work item (a multiple of 128)
};

// Compare: one (reference digest, target digest) pair after orientation.
struct CmpPair {
    uint64_t ref_first;    // first filter of the reference digest
    uint32_t ref_count;    // its filters
    uint32_t nch;          // chunks of the target digest
    uint64_t inst_begin;   // first instance index
    uint64_t partial_begin;// its first partial-maximum slot (nch per filter)
    uint32_t ref_set;      // 0: reference filters in set A, 1: in set B
    uint32_t pad;
};

// Text form: one digest of a set (sdbf::to_string layout).
struct TextDigest {
    uint64_t filt_first;   // first filter of the digest in the set
    uint64_t q_begin;      // first base64 quantum (3 input bytes) of the digest
    uint64_t body_off;     // text offset of the digest's first filter text
    uint64_t text_off;     // text offset of the digest's header
    uint32_t hdr_off;      // header bytes in the header blob
    uint32_t hdr_len;
    uint32_t bf_count;
    uint32_t dd;           // 0: stream mode (one base64 run), else block mode
};

// Text form, parsing: one digest's body in the uploaded text.
struct TextParse {
    uint64_t filt_first;   // destination filter of the digest in the set
    uint64_t q_begin;      // first base64 quantum of the digest
    uint64_t body_off;     // text offset of its base64 run (stream) or first ':' (block mode)
    uint32_t bf_count;
    uint32_t dd;           // 0: stream mode
    uint32_t max_elem;
    uint32_t last_count;
};

#endif
